"""GPU: the two-phase schedule (every game to a move budget, then only the parked survivors, in turns of that
budget from a work ring) changes nothing but the order of work: fitness, placements, work counters and trace
records equal the one-phase launch's, which equal the oracle's."""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O
import tetris_ai3_amd as T

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(ev, w, nm, seeds, cap):
    ev.stage(w, nm, seeds, cap)
    ev.launch()
    fit, pl = ev.fetch()
    return fit, pl, ev.fetch_counters()


@pytest.mark.parametrize("kind,wpg", [(O.MODEL_V11, 2), (O.MODEL_V11, 4), (O.MODEL_F16, 4), (O.MODEL_F32W, 4)])
def test_two_phases_equal_one_phase(kind, wpg):
    say(f"case: two phases kind={kind} wpg={wpg}")
    nm = 3
    w = np.random.default_rng(51).normal(0, 0.5 if kind != O.MODEL_F32W else 0.3,
                                         nm * O.num_params(kind)).astype(np.float32)
    seeds = O.expand_seeds(52, 6)
    ev = T.ModelEvaluator(model=kind)
    ev.set_launch(wpg, True, 0)
    ev.set_phases(0)
    one = run(ev, w, nm, seeds, 120)
    for budget in (1, 7, 40, 119, 120, 500, None):
        ev.set_phases(budget)
        two = run(ev, w, nm, seeds, 120)
        assert one[0].tobytes() == two[0].tobytes(), budget
        assert (one[1] == two[1]).all() and (one[2] == two[2]).all(), budget
    ev.close()
    if kind != O.MODEL_F32W:
        host, host_pl = O.eval_host_mt(kind, w, nm, seeds, 120)
        assert one[0].tobytes() == host.tobytes() and (one[1] == host_pl).all()


@pytest.mark.parametrize("wpg,grid", [(4, 2), (4, 3), (2, 1)])
def test_turns_on_few_workgroups(wpg, grid):
    """more parked games than workgroups: the second phase hands the games round in turns (many per game at a
    budget of 1-9 moves), re-parking each between turns; the results still equal the one-phase launch's"""
    say(f"case: turns wpg={wpg} grid={grid}")
    nm = 4
    w = np.random.default_rng(53).normal(0, 0.5, nm * O.num_params(O.MODEL_F16)).astype(np.float32)
    seeds = O.expand_seeds(54, 5)
    ev = T.ModelEvaluator(model=O.MODEL_F16)
    ev.set_launch(wpg, True, grid)
    ev.set_phases(0)
    one = run(ev, w, nm, seeds, 90)
    for budget in (1, 4, 9):
        ev.set_phases(budget)
        two = run(ev, w, nm, seeds, 90)
        assert one[0].tobytes() == two[0].tobytes(), budget
        assert (one[1] == two[1]).all() and (one[2] == two[2]).all(), budget
    ev.close()


def test_two_phase_traces_and_graph():
    say("case: two-phase traces and graph replay")
    kind, nm = O.MODEL_V11, 2
    w = np.random.default_rng(53).normal(0, 0.5, nm * 97).astype(np.float32)
    seeds = O.expand_seeds(54, 4)
    ev = T.ModelEvaluator(model=kind)
    ev.set_phases(0)
    cap = 4_000_000
    ev.set_trace(cap, every=3)
    ev.stage(w, nm, seeds, 50)
    ev.launch()
    ev.fetch()
    recs1, produced1 = ev.fetch_trace()
    assert 0 < produced1 <= cap  # every record kept, so the set cannot depend on the order of work
    ev.set_phases(10)
    ev.set_trace(cap, every=3)
    ev.stage(w, nm, seeds, 50)
    ev.launch()
    ev.fetch()
    recs2, produced2 = ev.fetch_trace()
    assert produced2 == produced1
    srt = lambda r: r[np.lexsort(r.T[::-1])]
    assert recs1.shape == recs2.shape and (srt(recs1) == srt(recs2)).all()
    ev.set_trace(0)
    ev.graph_prepare(nm, seeds, 50)
    fit_g, _ = ev.graph_run(w)
    fit_e = ev.eval(w, nm, seeds, 50)
    host, _ = O.eval_host_mt(kind, w, nm, seeds, 50)
    assert fit_g.tobytes() == host.tobytes() and fit_e.tobytes() == host.tobytes()
    ev.close()


def test_two_phase_pruned_and_autotune():
    say("case: two-phase pruned beam and autotune")
    kind, nm = O.MODEL_F16, 3
    w = np.random.default_rng(55).normal(0, 0.5, nm * O.num_params(kind)).astype(np.float32)
    seeds = O.expand_seeds(56, 5)
    ev = T.ModelEvaluator(model=kind)
    ev.set_pruner(T.pruner_init(3), 4.0, k=6, ply_mask=0b1110)
    ev.set_phases(0)
    one = run(ev, w, nm, seeds, 90)
    for budget in (5, 30):
        ev.set_phases(budget)
        two = run(ev, w, nm, seeds, 90)
        assert one[0].tobytes() == two[0].tobytes() and (one[1] == two[1]).all() and (one[2] == two[2]).all()
    ev.set_pruner(None, 0.0)
    T._TUNED.clear()
    tuned = T.autotune(ev, w, nm, seeds, 90, candidates=((2, 0), (4, 0)), phases=(0, 20))
    assert not tuned["cached"]
    assert set(tuned["timings_ms"]) == {"wpg2_grid0_phase0", "wpg2_grid0_phase20", "wpg4_grid0_phase0",
                                        "wpg4_grid0_phase20"}
    assert tuned["chosen"][0] in (2, 4) and tuned["chosen"][2] in (0, 20)
    host, host_pl = O.eval_host_mt(kind, w, nm, seeds, 90)
    fit, pl = run(ev, w, nm, seeds, 90)[:2]
    assert fit.tobytes() == host.tobytes() and (pl == host_pl).all()
    # the same batch shape again: the cached choice is applied without timing, on another evaluator too
    ev2 = T.ModelEvaluator(model=kind)
    again = T.autotune(ev2, w, nm, seeds, 90, candidates=((2, 0), (4, 0)), phases=(0, 20))
    assert again["cached"] and again["chosen"] == list(tuned["chosen"]) and again["key"] == tuned["key"]
    assert ev2.launch_info()["waves_per_game"] == tuned["chosen"][0]
    fit2, pl2 = run(ev2, w, nm, seeds, 90)[:2]
    assert fit2.tobytes() == host.tobytes()
    # a pruned evaluator tunes its own key (its kernel always runs 4 waves per game)
    ev2.set_pruner(T.pruner_init(3), 4.0, k=6, ply_mask=0b1110)
    pruned = T.autotune(ev2, w, nm, seeds, 90, phases=(0, 20))
    assert not pruned["cached"] and pruned["key"] != tuned["key"]
    assert set(pruned["timings_ms"]) == {"wpg4_grid0_phase0", "wpg4_grid0_phase20"}
    ev2.close()
    ev.close()


def test_auto_budget_parks_long_games():
    """the default first-phase budget (max_moves / 16 clamped to [128, 512]: 131 at a 2,100-piece cap) with games
    that outlive it: the reference's V11 champion, whose games run to the cap"""
    say("case: auto budget, V11 champion, 4 games x 2100 moves")
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "v11_champion.json")) as f:
        w = np.array(json.load(f)["weights_f64"], dtype=np.float64).astype(np.float32)
    seeds = O.expand_seeds(57, 4)
    ev = T.ModelEvaluator(model=O.MODEL_V11)
    ev.set_phases(0)
    one = run(ev, w, 1, seeds, 2100)
    ev.set_phases(None)
    auto = run(ev, w, 1, seeds, 2100)
    ev.close()
    assert (one[1] > 131).all()  # every game outlived the first phase
    assert one[0].tobytes() == auto[0].tobytes() and (one[1] == auto[1]).all() and (one[2] == auto[2]).all()
    host, host_pl = O.eval_host_mt(O.MODEL_V11, w, 1, seeds, 2100)
    assert auto[0].tobytes() == host.tobytes() and (auto[1] == host_pl).all()


def test_graph_prepared_first_with_two_phases():
    """graph_prepare as the first call of a fresh evaluator, with the two-phase schedule on: the parked-game
    buffers are allocated before the capture begins, and every replay equals the oracle"""
    say("case: graph prepared first, two phases")
    kind, nm, cap = O.MODEL_F16, 2, 60
    w = np.random.default_rng(56).normal(0, 0.5, nm * O.num_params(kind)).astype(np.float32)
    seeds = O.expand_seeds(57, 3)
    ev = T.ModelEvaluator(model=kind)
    ev.set_phases(15)
    ev.graph_prepare(nm, seeds, cap)
    host, host_pl = O.eval_host_mt(kind, w, nm, seeds, cap)
    assert (host_pl > 15).any()  # some game is parked and resumed
    for rep in range(2):
        fit, pl = ev.graph_run(w)
        assert fit.tobytes() == host.tobytes() and (pl == host_pl).all(), rep
    ev.close()
