"""GPU: narrower beams (runtime widths at most the reference's 16 / 32 / 32, variation_sequences.hpp:22) play the
games of the oracle's search built with those widths, bit for bit: fitness, placements and work counters. These
are the non-learned way to spend fewer value evaluations, the baseline the learned pruner is measured against."""
import sys

import numpy as np
import pytest

import oracle as O
import tetris_ai3_amd as T

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", [T.MODEL_V11, T.MODEL_F16])
def test_narrow_beams_equal_oracle(kind):
    nm, cap = 3, 150
    P = T.num_params(kind)
    w = np.random.default_rng(120 + kind).normal(0, 0.5, nm * P).astype(np.float32)
    seeds = T.expand_seeds(121, 4)
    ev = T.ModelEvaluator(model=kind)
    for ks in ((8, 16, 16), (4, 8, 8), (2, 4, 4), (1, 1, 1), (16, 8, 32), (16, 32, 32)):
        say(f"case: beam widths {ks} kind={kind}")
        ev.set_beam_widths(ks)
        ev.stage(w, nm, seeds, cap)
        ev.launch()
        fit, pl = ev.fetch()
        cnt = ev.fetch_counters()
        c = O.Counters()
        with O.widths(ks):
            want, want_pl = O.eval_host_mt(kind, w, nm, seeds, cap, counters=c)
        assert fit.tobytes() == want.tobytes(), ks
        assert np.array_equal(pl, want_pl), ks
        # the device scores a repeated list's content once and takes the roots the previous move scored (the
        # oracle counts both)
        assert int(cnt[:, 0].sum()) == c.applies - c.repeat_applies - c.reuse_applies, ks
        assert int(cnt[:, 1].sum()) == c.legal - c.repeat_legal - c.reuse_legal, ks
    ev.close()


def test_narrow_beam_with_pruner_traces_and_graph():
    """the widths hold in the pruned kernel (k = 34 is the full beam) and in a generation graph"""
    kind, nm, cap, ks = T.MODEL_F16, 2, 80, (6, 12, 10)
    w = np.random.default_rng(122).normal(0, 0.5, nm * T.num_params(kind)).astype(np.float32)
    seeds = T.expand_seeds(123, 3)
    with O.widths(ks):
        want, want_pl = O.eval_host_mt(kind, w, nm, seeds, cap)
    ev = T.ModelEvaluator(model=kind)
    ev.set_beam_widths(ks)
    say("case: narrow beam, pruned k=34")
    ev.set_pruner(T.pruner_init(124), 4.0, k=34, ply_mask=0b1110)
    assert ev.eval(w, nm, seeds, cap).tobytes() == want.tobytes()
    ev.set_pruner(None, 0.0)
    say("case: narrow beam, graph")
    ev.graph_prepare(nm, seeds, cap)
    for _ in range(2):
        fit, pl = ev.graph_run(w)
        assert fit.tobytes() == want.tobytes() and np.array_equal(pl, want_pl)
    ev.close()
