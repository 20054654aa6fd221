"""GPU: one CMA-ES generation as a captured HIP graph.

Every replay must equal the eager evaluation of the same weights bit for bit. Replays are interleaved with
eager work on another evaluator and host synchronisations, and each one is checked on its own.
"""
import sys

import numpy as np
import pytest

import oracle as O
import tetris_ai3_amd as T

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", [O.MODEL_V11, O.MODEL_F16])
def test_graph_replays_equal_eager(kind):
    say(f"case: graph replays kind={kind}")
    nm, cap = 3, 70
    seeds = O.expand_seeds(41, 4)
    P = O.num_params(kind)
    rng = np.random.default_rng(42)
    gr = T.ModelEvaluator(model=kind)
    gr.graph_prepare(nm, seeds, cap)
    eager = T.ModelEvaluator(model=kind)
    for rep in range(3):
        w = rng.normal(0, 0.5, nm * P).astype(np.float32)
        fit, pl = gr.graph_run(w)
        assert gr.last_kernel_ms() > 0
        other = rng.normal(0, 0.5, 2 * P).astype(np.float32)
        eager.eval(other, 2, seeds[:2], 30)  # unrelated eager work between replays
        eager.stage(w, nm, seeds, cap)
        eager.launch()
        want, want_pl = eager.fetch()
        assert fit.tobytes() == want.tobytes() and (pl == want_pl).all(), rep
        if rep == 0:
            host, host_pl = O.eval_host_mt(kind, w, nm, seeds, cap)
            assert fit.tobytes() == host.tobytes() and (pl == host_pl).all()
    gr.close()
    eager.close()


def test_pruned_graph_equals_eager_pruned():
    say("case: pruned graph")
    kind, nm, cap = O.MODEL_F16, 2, 60
    seeds = O.expand_seeds(43, 3)
    w = np.random.default_rng(44).normal(0, 0.5, nm * O.num_params(kind)).astype(np.float32)
    pw = T.pruner_init(45)
    gr = T.ModelEvaluator(model=kind)
    gr.set_pruner(pw, 5.0, k=6, ply_mask=0b1110)
    gr.graph_prepare(nm, seeds, cap)
    fit1, _ = gr.graph_run(w)
    fit2, _ = gr.graph_run(w)
    want, _, _ = O.eval_host_pruned_mt(kind, w, nm, seeds, cap, pw, 5.0, 6, 0b1110)
    assert fit1.tobytes() == want.tobytes() and fit2.tobytes() == want.tobytes()
    gr.close()


def test_stage_discards_the_graph():
    say("case: stage discards graph")
    ev = T.ModelEvaluator()
    seeds = O.expand_seeds(46, 2)
    ev.graph_prepare(1, seeds, 10)
    ev.eval(np.zeros(97, dtype=np.float32), 1, seeds, 10)
    with pytest.raises(T.HipError, match="no generation graph"):
        ev.graph_run(np.zeros(97, dtype=np.float32))
    ev.close()
