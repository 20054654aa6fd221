"""GPU: the per-item move log against the CPU search, move by move.

- Bit-exact networks (V11, 16-32-32-1): every logged move of every game equals the oracle's own game record
  (hold, rotation, x of each placement, eval_host's loop, model_eval_host.cpp:37-47), with one and two phases,
  with garbage rows, and from a generation graph.
- The wide network (BF16 MFMA against the oracle's FP32 restatement): every device game is replayed on the
  oracle's engine and, at each state the device reached, the oracle's search picks its own move. The fraction
  of states where both pick the same placement is the placement-agreement rate; the bound in the test is the
  tolerance the BF16 rounding leaves.
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
import tetris_ai3_amd as T

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def encode(moves):
    """oracle game record rows (hold, o, x + 2, cleared) -> the move-log bytes"""
    m = moves.astype(np.uint32)
    return ((m[:, 0] << 7) | (m[:, 1] << 5) | m[:, 2]).astype(np.uint8)


def oracle_logs(kind, w, nm, seeds, cap):
    P = O.num_params(kind)
    out = []
    for m in range(nm):
        for s in seeds:
            moves, _, fit = O.play_trace(kind, w[m * P:(m + 1) * P], int(s), cap)
            out.append((encode(moves), np.float32(fit)))
    return out


@pytest.mark.parametrize("kind", [O.MODEL_V11, O.MODEL_F16])
def test_move_log_equals_oracle_games(kind):
    nm, cap = 3, 90
    seeds = O.expand_seeds(61 + kind, 4)
    w = np.random.default_rng(62 + kind).normal(0, 0.5, nm * O.num_params(kind)).astype(np.float32)
    want = oracle_logs(kind, w, nm, seeds, cap)
    ev = T.ModelEvaluator(model=kind)
    ev.set_move_log(True)
    for phases in (0, 17):
        say(f"case: move log kind={kind} phases={phases}")
        ev.set_phases(phases)
        ev.stage(w, nm, seeds, cap)
        ev.launch()
        fit, pl = ev.fetch()
        log = ev.fetch_moves()
        assert log.shape == (nm * len(seeds), cap)
        for i, (wm, wf) in enumerate(want):
            assert pl[i] == wm.size and fit[i] == wf, (phases, i)
            assert (log[i, :pl[i]] == wm).all(), (phases, i, log[i, :pl[i]], wm)
    ev.close()


def test_move_log_garbage_and_graph():
    kind, nm, cap = O.MODEL_F16, 2, 70
    seeds = O.expand_seeds(63, 3)
    w = np.random.default_rng(64).normal(0, 0.5, nm * O.num_params(kind)).astype(np.float32)
    with O.garbage():
        want = oracle_logs(kind, w, nm, seeds, cap)
    ev = T.ModelEvaluator(model=kind)
    ev.set_garbage(True)
    ev.set_move_log(True)
    ev.set_phases(20)
    say("case: move log with garbage rows, graph replay")
    ev.graph_prepare(nm, seeds, cap)
    for rep in range(2):
        fit, pl = ev.graph_run(w)
        log = ev.fetch_moves()
        for i, (wm, wf) in enumerate(want):
            assert pl[i] == wm.size and fit[i] == wf and (log[i, :pl[i]] == wm).all(), (rep, i)
    ev.close()


def test_move_log_with_traces_and_pruned_k34():
    """the log is kept alongside traces, and by the pruned kernel (k = 34 is the full beam)"""
    kind, nm, cap = O.MODEL_F16, 2, 60
    seeds = O.expand_seeds(67, 3)
    w = np.random.default_rng(68).normal(0, 0.5, nm * O.num_params(kind)).astype(np.float32)
    want = oracle_logs(kind, w, nm, seeds, cap)
    ev = T.ModelEvaluator(model=kind)
    ev.set_move_log(True)
    for case in ("traces", "pruned k=34"):
        say(f"case: move log with {case}")
        if case == "traces":
            ev.set_trace(2_000_000, every=4)
        else:
            ev.set_pruner(T.pruner_init(69), 4.0, k=34, ply_mask=0b1110)
        ev.stage(w, nm, seeds, cap)
        ev.launch()
        fit, pl = ev.fetch()
        log = ev.fetch_moves()
        for i, (wm, wf) in enumerate(want):
            assert pl[i] == wm.size and fit[i] == wf and (log[i, :pl[i]] == wm).all(), (case, i)
        if case == "traces":
            assert ev.fetch_trace()[1] > 0
    ev.close()


def test_move_log_off_is_a_state_error():
    ev = T.ModelEvaluator()
    say("case: move log not enabled")
    ev.eval(np.zeros(97, dtype=np.float32), [1], 5)
    with pytest.raises(T.HipError, match="move log"):
        ev.fetch_moves()
    ev.close()


def replay_agreement(kind, w_model, seed, log_row, n):
    """replays a device game on the oracle engine; returns the number of states where the oracle's search picks
    the device's move"""
    e = O.engine(int(seed))
    same = 0
    for i in range(n):
        b = int(log_row[i])
        hold, o, x = b >> 7, (b >> 5) & 3, (b & 15) - 2
        r = O.search_beam(e, kind, w_model)
        if r is not None and (int(r["hold"]), r["o"], r["x"]) == (hold, o, x):
            same += 1
        if hold:
            O.engine_hold(e)
        if O.engine_place(e, o, x) < 0:
            break
    return same


def test_wide_placement_agreement():
    """32-128-128-1 on BF16 MFMA: per-move agreement with the FP32 oracle along the device's own games (with
    garbage rows, as the scaled workload plays). Measured on the MI355X: 649 of 649 moves agree."""
    kind, nm, cap = O.MODEL_F32W, 4, 40
    P = O.num_params(kind)
    seeds = O.expand_seeds(65, 6)
    w = np.random.default_rng(66).normal(0, 0.3, nm * P).astype(np.float32)
    ev = T.ModelEvaluator(model=kind)
    ev.set_garbage(True)
    ev.set_move_log(True)
    say("case: wide move log")
    ev.stage(w, nm, seeds, cap)
    ev.launch()
    _, pl = ev.fetch()
    log = ev.fetch_moves()
    ev.close()
    items = [(m, g) for m in range(nm) for g in range(len(seeds))]
    say(f"case: wide replay of {int(pl.sum())} device moves on the oracle")
    with O.garbage(), ThreadPoolExecutor(16) as pool:
        same = list(pool.map(lambda it: replay_agreement(kind, w[it[0] * P:(it[0] + 1) * P], seeds[it[1]],
                                                         log[it[0] * len(seeds) + it[1]],
                                                         int(pl[it[0] * len(seeds) + it[1]])), items))
    rate = sum(same) / float(pl.sum())
    say(f"wide placement agreement {rate:.4f} over {int(pl.sum())} moves")
    assert rate >= 0.99


def test_preview_boards_equal_oracle_games(tmp_path):
    """the preview (tetris_ai3_amd.preview): games played on the GPU with the move log, boards rebuilt by the host
    replay, equal to the oracle's own games board by board; the command writes its frames"""
    import json
    import os
    from tetris_ai3_amd import preview
    here = os.path.dirname(os.path.abspath(__file__))
    w = np.array(json.load(open(os.path.join(here, "golden", "v11_champion.json")))["weights_f64"], dtype=np.float32)
    seeds = O.expand_seeds(71, 3)
    say("case: preview, v11 champion 3 games x 120 moves")
    ev = T.ModelEvaluator(model=O.MODEL_V11)
    games = preview.play(ev, w, 1, seeds, 120, False)
    ev.close()
    for g, s in zip(games, seeds):
        moves, boards, fit = O.play_trace(O.MODEL_V11, w, int(s), 120)
        assert g["placements"] == moves.shape[0] and np.float32(g["fitness"]) == np.float32(fit)
        assert np.array_equal(g["boards"], boards)
    champ = tmp_path / "v11.champ"
    T.champ_write(str(champ), w.astype(np.float64))
    out = tmp_path / "frames.txt"
    assert preview.main(["--champ", str(champ), "--games", "2", "--moves", "30", "--seed", "5", "--every", "10",
                         "--out", str(out)]) == 0
    text = out.read_text()
    assert text.count("game 0") >= 4 and "fitness" in text
