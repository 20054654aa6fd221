"""GPU: two evaluators used from two host threads at once on one GPU (the reference pools one evaluator per
scheduler thread, two by default: ta3_trainer/runner/src/main.cpp:27; tetris_problem2.hpp:27). Each has its own
stream, queue and buffers, so their kernels may overlap; every result must equal the oracle's."""
import sys
import threading

import numpy as np
import pytest

import oracle as O
import tetris_ai3_amd as T

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_two_evaluators_in_two_threads():
    say("case: two evaluators, two threads, V11 and F16")
    jobs = []
    for kind, seed in ((O.MODEL_V11, 91), (O.MODEL_F16, 92)):
        nm = 6
        w = np.random.default_rng(seed).normal(0, 0.5, nm * O.num_params(kind)).astype(np.float32)
        seeds = O.expand_seeds(seed, 5)
        want, want_pl = O.eval_host_mt(kind, w, nm, seeds, 120)
        jobs.append((kind, w, nm, seeds, want, want_pl))
    evs = [T.ModelEvaluator(model=j[0]) for j in jobs]
    results = [[] for _ in jobs]
    errors = []
    start = threading.Barrier(len(jobs))

    def work(i):
        kind, w, nm, seeds, _, _ = jobs[i]
        try:
            start.wait()
            for _ in range(3):
                evs[i].stage(w, nm, seeds, 120)
                evs[i].launch()
                results[i].append(evs[i].fetch())
        except Exception as exc:  # reported by the main thread
            errors.append(repr(exc))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    for ev in evs:
        ev.close()
    assert not errors, errors
    assert not any(t.is_alive() for t in threads)
    for (kind, _, _, _, want, want_pl), res in zip(jobs, results):
        assert len(res) == 3
        for fit, pl in res:
            assert fit.tobytes() == want.tobytes() and (pl == want_pl).all(), kind
