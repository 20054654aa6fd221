"""GPU: the learned pruner's traces, training kernels, inference and pruned beam against the CPU oracle.

- Traces of the full beam: the device's records equal the oracle's as multisets (bit-exact boards, metas and
  labels; the device appends them in no particular order).
- Training: the device gradient (FP32, bfloat16 forward roundings) against the oracle's float64 gradient with
  the same roundings, within a stated tolerance.
- Inference: integer grids on BF16 MFMA, so device logits equal the oracle's bit for bit.
- Pruned beam: k = 34 is the full beam; smaller k is bit-exact against the oracle's pruned beam.
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


def models(kind, n, seed):
    rng = np.random.default_rng(seed)
    P = O.num_params(kind)
    return rng.normal(0, 0.5, n * P).astype(np.float32)


def oracle_traces(kind, w, num_models, seeds, max_moves, every):
    P = O.num_params(kind)
    out = []
    for m in range(num_models):
        for s in seeds:
            out.append(O.trace_records(kind, w[m * P:(m + 1) * P], int(s), max_moves, every))
    return np.concatenate(out)


def as_sorted_rows(recs):
    r = np.ascontiguousarray(recs, dtype=np.uint32).reshape(-1, 11)
    return r[np.lexsort(r.T[::-1])]


@pytest.fixture(scope="module")
def trace_set():
    """device traces of a small V11 batch (checked against the oracle in the first test)"""
    kind, nm = O.MODEL_V11, 2
    w = models(kind, nm, 1)
    seeds = O.expand_seeds(31, 2)
    ev = T.ModelEvaluator(model=kind)
    ev.set_trace(200000, every=5)
    ev.stage(w, nm, seeds, 40)
    ev.launch()
    fit, pl = ev.fetch()
    recs, produced = ev.fetch_trace()
    ev.close()
    return dict(kind=kind, w=w, nm=nm, seeds=seeds, fit=fit, pl=pl, recs=recs, produced=produced)


@pytest.mark.parametrize("kind", [O.MODEL_V11, O.MODEL_F16])
def test_traces_match_oracle(kind, trace_set):
    if kind == O.MODEL_V11:
        t = trace_set
        w, nm, seeds, recs, produced, fit, pl = (t["w"], t["nm"], t["seeds"], t["recs"], t["produced"], t["fit"],
                                                 t["pl"])
        cap = 40
    else:
        say("case: traces F16 2 models x 2 seeds, 30 moves, every 4")
        nm, cap = 2, 30
        w = models(kind, nm, 2)
        seeds = O.expand_seeds(32, 2)
        ev = T.ModelEvaluator(model=kind)
        ev.set_trace(200000, every=4)
        ev.stage(w, nm, seeds, cap)
        ev.launch()
        fit, pl = ev.fetch()
        recs, produced = ev.fetch_trace()
        ev.close()
    every = 5 if kind == O.MODEL_V11 else 4
    want = oracle_traces(kind, w, nm, seeds, cap, every)
    assert produced == want.shape[0] == recs.shape[0]
    assert (as_sorted_rows(recs) == as_sorted_rows(want)).all()
    # the traced launch still plays the full beam
    host, host_pl = O.eval_host_mt(kind, w, nm, seeds, cap)
    assert fit.tobytes() == host.tobytes() and (pl == host_pl).all()


def test_trace_capacity_drops_the_excess():
    say("case: trace cap 1000")
    w = models(O.MODEL_V11, 1, 3)
    seeds = O.expand_seeds(33, 2)
    ev = T.ModelEvaluator(model=O.MODEL_V11)
    ev.set_trace(1000, every=1)
    ev.stage(w, 1, seeds, 20)
    ev.launch()
    fit, _ = ev.fetch()
    recs, produced = ev.fetch_trace()
    ev.close()
    full = oracle_traces(O.MODEL_V11, w, 1, seeds, 20, 1)
    assert produced == full.shape[0] > 1000 and recs.shape[0] == 1000
    # every kept record is one of the oracle's
    keys = {r.tobytes() for r in full}
    assert all(r.tobytes() in keys for r in recs)
    assert fit.tobytes() == O.eval_host_mt(O.MODEL_V11, w, 1, seeds, 20)[0].tobytes()


def test_gradient_matches_oracle(trace_set):
    """device FP32 gradient vs the oracle's float64 gradient (same bfloat16 roundings): relative L2 error
    <= 1e-4 per parameter block, loss within 1e-5 relative"""
    say("case: pruner gradient 600 samples")
    recs = trace_set["recs"]
    rng = np.random.default_rng(4)
    smp = recs[rng.choice(recs.shape[0], 600, replace=False)]
    w = T.pruner_init(5) * np.float32(0.5)
    w[64 * 265:64 * 266] = rng.normal(0, 0.1, 64)  # nonzero biases
    ev = T.ModelEvaluator()
    loss, g = ev.pruner_grad(w, smp, pos_weight=3.0)
    ev.close()
    want_loss, want = O.pruner_loss_grad(w, smp, pos_weight=3.0, rounded=True)
    assert abs(loss - want_loss) <= 1e-5 * abs(want_loss)
    blocks = [(0, 64 * 265), (64 * 265, 64 * 266), (64 * 266, 64 * 266 + 4096), (64 * 266 + 4096, 64 * 266 + 4160),
              (64 * 266 + 4160, 21248), (21248, 21249)]
    for lo, hi in blocks:
        num = np.linalg.norm(g[lo:hi].astype(np.float64) - want[lo:hi])
        den = np.linalg.norm(want[lo:hi])
        assert num <= 1e-4 * den + 1e-9, (lo, hi, num, den)


def test_device_logits_equal_oracle(trace_set):
    say("case: pruner logits 3000 samples")
    recs = trace_set["recs"]
    rng = np.random.default_rng(6)
    smp = recs[rng.choice(recs.shape[0], 3000, replace=False)]
    w = T.pruner_init(7)
    ev = T.ModelEvaluator()
    for act in (0.75, 3.0, 40.0):
        got = ev.pruner_logits(w, act, smp)
        want = O.pruner_scores(w, act, smp)
        assert got.tobytes() == want.tobytes(), (act, np.nonzero(got != want)[0][:5])
    ev.close()


@pytest.mark.parametrize("kind", [O.MODEL_V11, O.MODEL_F16])
def test_pruned_k34_is_the_full_beam(kind):
    say(f"case: pruned k=34 kind={kind}")
    nm = 2
    w = models(kind, nm, 8)
    seeds = O.expand_seeds(34, 3)
    ev = T.ModelEvaluator(model=kind)
    full = ev.eval(w, nm, seeds, 60)
    ev.set_pruner(T.pruner_init(9), 4.0, k=34, ply_mask=0b1110)
    got = ev.eval(w, nm, seeds, 60)
    ev.close()
    assert got.tobytes() == full.tobytes()
    assert got.tobytes() == O.eval_host_mt(kind, w, nm, seeds, 60)[0].tobytes()


@pytest.mark.parametrize("kind,k,mask", [(O.MODEL_V11, 8, 0b1100), (O.MODEL_V11, 3, 0b1110),
                                         (O.MODEL_F16, 8, 0b1100), (O.MODEL_F16, 5, 0b1010)])
def test_pruned_beam_matches_oracle(kind, k, mask):
    """bit-exact fitness, placements and work counters against the oracle's pruned beam"""
    say(f"case: pruned kind={kind} k={k} mask={mask:#b}")
    nm = 2
    w = models(kind, nm, 10 + k)
    seeds = O.expand_seeds(35, 3)
    pw = T.pruner_init(11)
    act = 6.0
    ev = T.ModelEvaluator(model=kind)
    ev.set_pruner(pw, act, k=k, ply_mask=mask)
    ev.stage(w, nm, seeds, 80)
    ev.launch()
    fit, pl = ev.fetch()
    cnt = ev.fetch_counters()
    ev.close()
    want, want_pl, want_cnt = O.eval_host_pruned_mt(kind, w, nm, seeds, 80, pw, act, k, mask)
    assert fit.tobytes() == want.tobytes()
    assert (pl == want_pl).all()
    # repeated lists (same parent board, clear histogram, piece and held-is-I as an earlier list of the level) are
    # scored once on the device, pruner included, and the roots the previous move scored are taken from it (V11 and
    # 16-32-32-1, when level 1 is not pruned); the oracle counts both
    reuse_a, reuse_l = want_cnt.reuse_applies, want_cnt.reuse_legal
    assert (reuse_l > 0) == ((mask & 0b10) == 0)
    assert int(cnt[:, 0].sum()) == want_cnt.applies - want_cnt.repeat_applies - reuse_a
    assert int(cnt[:, 1].sum()) == want_cnt.legal - want_cnt.repeat_legal - reuse_l


def test_training_learns_the_labels(trace_set):
    """a few device epochs on the traces: the loss falls and the trained logits separate the labels
    (ROC AUC on held-out records) far better than the initial network"""
    say("case: pruner training")
    from tetris_ai3_amd import learn
    recs = learn.canonical_order(trace_set["recs"])
    rng = np.random.default_rng(12)
    perm = rng.permutation(recs.shape[0])
    n_train = int(0.8 * len(perm))
    tr, te = recs[perm[:n_train]], recs[perm[n_train:]]
    w0 = T.pruner_init(13)
    ev = T.ModelEvaluator()
    w1, losses = ev.pruner_train(w0, tr, epochs=4, batch=256, lr=2e-3, pos_weight=2.0, seed=1)
    assert losses[-1] < 0.8 * losses[0], losses

    def auc(w):
        h = ev.pruner_hidden(w, te)
        act = float(np.quantile(h[h > 0], 0.9999)) if (h > 0).any() else 1.0
        s = ev.pruner_logits(w, act, te).astype(np.float64)
        y = (te[:, 10] >> 11) & 1
        pos, neg = s[y == 1], s[y == 0]
        order = np.argsort(np.concatenate([pos, neg]), kind="stable")
        ranks = np.empty(len(order))
        ranks[order] = np.arange(1, len(order) + 1)
        return (ranks[:len(pos)].sum() - len(pos) * (len(pos) + 1) / 2) / (len(pos) * len(neg))

    a0, a1 = auc(w0), auc(w1)
    ev.close()
    say(f"auc untrained {a0:.3f} trained {a1:.3f}")
    assert a1 > 0.8 and a1 > a0 + 0.15


def test_trained_pruner_gain_over_untrained():
    """What the learned part gains over the method without it. The V11 champion plays 32 games with the full
    beam, whose traces (1 move in 8) train the pruner on the device. On 48 other games the pruned beam (k = 4 at
    plies 2 and 3) with the trained network keeps much more of the full beam's mean score than the same pruned
    beam ranked by the untrained network, at an equal cut in value evaluations. Measured on the MI355X: full beam
    294.25 with 3,509 value evaluations per placement; trained pruner 290.67 (1.22% loss) with 1,262; untrained
    155.72 (47% loss) with 1,241."""
    import json
    import os
    from tetris_ai3_amd import learn
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "v11_champion.json")) as f:
        w = np.array(json.load(f)["weights_f64"], dtype=np.float64).astype(np.float32)
    cap = 300
    ev = T.ModelEvaluator(model=O.MODEL_V11)
    say("case: pruner gain, traces of 32 champion games")
    recs, produced, _ = learn.collect_traces(ev, w, 1, O.expand_seeds(901, 32), cap, every=8, cap=8_000_000)
    assert recs.shape[0] == produced > 100_000
    say(f"case: pruner gain, training on {recs.shape[0]} records")
    pw, losses, _ = learn.train_pruner(ev, recs, epochs=4, lr=2e-3, seed=3)
    pw0 = T.pruner_init(3)
    seeds = O.expand_seeds(902, 48)

    def run():
        ev.stage(w, 1, seeds, cap)
        ev.launch()
        fit, pl = ev.fetch()
        cnt = ev.fetch_counters()
        return float(fit.astype(np.float64).mean()), float(cnt[:, 1].sum()) / float(pl.sum())

    say("case: pruner gain, full beam")
    ev.set_pruner(None, 0.0)
    full, full_evals = run()
    res = {}
    for name, p in (("untrained", pw0), ("trained", pw)):
        say(f"case: pruner gain, {name} pruner k=4 plies 2-3")
        ev.set_pruner(p, learn.calibrate_act_range(ev, p, recs), k=4, ply_mask=0b1100)
        res[name] = run()
    ev.close()
    loss = {n: 1.0 - s / full for n, (s, _) in res.items()}
    say(f"full {full:.2f} ({full_evals:.0f} evals/placement) "
        + " ".join(f"{n} {s:.2f} loss {100 * loss[n]:.2f}% ({e:.0f} evals/placement)" for n, (s, e) in res.items())
        + f" epoch losses {[round(float(x), 3) for x in losses]}")
    assert losses[-1] < losses[0]
    for n, (_, e) in res.items():
        assert e < 0.45 * full_evals, (n, e, full_evals)
    assert loss["trained"] < 0.02 and loss["untrained"] > 0.2, loss
