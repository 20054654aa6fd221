"""GPU: the benchmark workloads themselves, at full size.

- The headline generation (4 islands x 50 population x 50 games = 10,000 simulations, cap 2,000 pieces, the
  16-32-32-1 network, CMA-ES-sampled weights: bench.py's step): size-independent properties of the whole batch
  (counters, placements, determinism, lane utilisation) and bit-exact fitness, placements and work counters
  against the oracle for the first 16 items and the 8 longest games.
- The scaled generation (256 models x 100 games = 25,600 simulations, cap 5,000, garbage rows, the 32-128-128-1
  network on BF16 MFMA): properties of the whole batch, then a slice of its longest and first games re-played with
  the move log (the trace kernel, which holds its weights in device memory instead of LDS) must reproduce the
  same games, and along those games the oracle's FP32 search picks the device's move at >= 99% of the states.
- The pruned beam on the wide network at k < 34, move by move against the oracle's pruned search.
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import bench
import oracle as O
import tetris_ai3_amd as T

pytestmark = pytest.mark.gpu

MAX_APPLIES_PER_MOVE = 68 + 16 * 2 * 34 + 32 * 2 * 34 * 2  # roots + level 1 + level 2 + leaves


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check_batch_properties(fit, pl, cnt, cap):
    assert np.isfinite(fit).all()
    assert (pl >= 1).all() and (pl <= cap).all()
    applies, legal, lanes = cnt[:, 0], cnt[:, 1], cnt[:, 2]
    assert (legal >= pl).all() and (applies >= legal).all() and (lanes >= applies).all()
    assert (lanes % 64 == 0).all()
    assert (applies <= MAX_APPLIES_PER_MOVE * pl.astype(np.uint64)).all()
    util = float(legal.sum()) / float(lanes.sum())
    assert 0.9 < util <= 1.0
    return util


def replay_agreement(kind, w_model, seed, log_row, n, search):
    """replays a device game on the oracle's engine; counts the states where the oracle's search picks the
    device's move"""
    e = O.engine(int(seed))
    same = 0
    for i in range(n):
        b = int(log_row[i])
        hold, o, x = b >> 7, (b >> 5) & 3, (b & 15) - 2
        r = search(e, kind, w_model)
        if r is not None and (int(r["hold"]), r["o"], r["x"]) == (hold, o, x):
            same += 1
        if hold:
            O.engine_hold(e)
        if O.engine_place(e, o, x) < 0:
            break
    return same


def test_config1_generation_bit_exact_sample_and_properties():
    kind = T.MODEL_F16
    w, M, seeds = bench.workload(kind)
    G, cap, P = len(seeds), bench.CAP, T.num_params(kind)
    ev = T.ModelEvaluator(model=kind)
    say(f"case: config1 generation {M} x {G}, cap {cap}")
    ev.stage(w, M, seeds, cap)
    ev.launch()
    fit, pl = ev.fetch()
    cnt = ev.fetch_counters()
    ev.launch()
    fit2, pl2 = ev.fetch()
    # the move-log (trace) kernel scores every list and every root stage, as the reference does; the plain
    # kernel skips repeated lists and reuses roots: every game of the generation must come out the same (the
    # agreement reports compare the move-log kernel with the oracle move by move over all 10,000 games)
    say("case: config1 generation, plain kernel vs the kernel that scores every candidate")
    ev.set_move_log(True)
    ev.stage(w, M, seeds, cap)
    ev.launch()
    fit3, pl3 = ev.fetch()
    cnt3 = ev.fetch_counters()
    ev.close()
    assert fit.tobytes() == fit2.tobytes() and np.array_equal(pl, pl2)  # a relaunch repeats bit for bit
    assert fit.tobytes() == fit3.tobytes() and np.array_equal(pl, pl3)
    skipped = 1.0 - cnt[:, 1].sum() / cnt3[:, 1].sum()
    say(f"value evaluations skipped by the plain kernel: {skipped:.4f}")
    assert 0.15 < skipped < 0.35
    util = check_batch_properties(fit, pl, cnt, cap)
    assert (pl == cap).sum() >= 100  # good models reach the cap (uneven lengths)
    say(f"placements {int(pl.sum())}, p50 {np.percentile(pl, 50)}, at cap {(pl == cap).sum()}, lane util {util:.4f}")

    longest = np.argsort(-pl.astype(np.int64), kind="stable")[:8]
    items = sorted(set(range(16)) | set(int(i) for i in longest))

    def oracle_item(i):
        m, g = divmod(i, G)
        c = O.Counters()
        f, p = O.eval_host_mt(kind, w[m * P:(m + 1) * P], 1, seeds[g:g + 1], cap, threads=1, counters=c)
        # repeated lists are scored once, the roots the previous move scored are taken from it
        return f[0], p[0], c.applies - c.repeat_applies - c.reuse_applies, c.legal - c.repeat_legal - c.reuse_legal

    say(f"case: config1 oracle on {len(items)} items ({int(pl[items].sum())} placements)")
    with ThreadPoolExecutor(16) as pool:
        want = list(pool.map(oracle_item, items))
    for i, (f, p, applies, legal) in zip(items, want):
        assert np.float32(f).tobytes() == fit[i].tobytes(), (i, f, fit[i])
        assert p == pl[i] and applies == cnt[i, 0] and legal == cnt[i, 1], i


def test_config4_scaled_generation_and_wide_agreement():
    kind = T.MODEL_F32W
    w, M, seeds = bench.wide_workload()
    G, cap, P = len(seeds), bench.WIDE_CAP, T.num_params(kind)
    ev = T.ModelEvaluator(model=kind)
    ev.set_garbage(True)
    say(f"case: config4 generation {M} x {G}, cap {cap}, garbage rows")
    ev.stage(w, M, seeds, cap)
    ev.launch()
    fit, pl = ev.fetch()
    cnt = ev.fetch_counters()
    ev.close()
    util = check_batch_properties(fit, pl, cnt, cap)
    say(f"placements {int(pl.sum())}, p50 {np.percentile(pl, 50)}, p99 {np.percentile(pl, 99)}, util {util:.4f}")

    # the slice: the 4 longest games and the first 4 items, re-played by the trace kernel with the move log
    chosen = list(np.argsort(-pl.astype(np.int64), kind="stable")[:4]) + [0, 1, 2, 3]
    models = sorted(set(int(i) // G for i in chosen))
    games = sorted(set(int(i) % G for i in chosen))
    ws = np.concatenate([w[m * P:(m + 1) * P] for m in models])
    ev = T.ModelEvaluator(model=kind)
    ev.set_garbage(True)
    ev.set_move_log(True)
    say(f"case: config4 slice {len(models)} models x {len(games)} games with the move log")
    ev.stage(ws, len(models), seeds[games], cap)
    ev.launch()
    sfit, spl = ev.fetch()
    log = ev.fetch_moves()
    ev.close()
    rows = []
    for i in chosen:
        m, g = divmod(int(i), G)
        r = models.index(m) * len(games) + games.index(g)
        assert sfit[r].tobytes() == fit[i].tobytes() and spl[r] == pl[i], (i, sfit[r], fit[i])
        rows.append((m, g, r))

    n_moves = 120
    say(f"case: config4 oracle replay of {len(rows)} games x first {n_moves} moves")
    with O.garbage(), ThreadPoolExecutor(16) as pool:
        same = list(pool.map(lambda t: replay_agreement(kind, w[t[0] * P:(t[0] + 1) * P], seeds[t[1]], log[t[2]],
                                                        min(n_moves, int(spl[t[2]])), O.search_beam), rows))
    total = sum(min(n_moves, int(spl[r])) for _, _, r in rows)
    rate = sum(same) / float(total)
    say(f"wide placement agreement {rate:.4f} over {total} moves")
    assert rate >= 0.99


@pytest.mark.parametrize("k,mask", [(8, 0b1000), (6, 0b1100)])
def test_wide_pruned_beam_moves_agree_with_oracle_pruned_search(k, mask):
    """k < 34: the device's pruned wide beam (exact integer-grid pruner, BF16 MFMA value network) and the oracle's
    pruned search (same pruner logits, FP32 restatement of the network) pick the same move at >= 99% of the states
    along the device's games"""
    kind, nm, cap = T.MODEL_F32W, 3, 40
    P = T.num_params(kind)
    w = np.random.default_rng(81).normal(0, 0.3, nm * P).astype(np.float32)
    seeds = T.expand_seeds(82, 4)
    pw, act = T.pruner_init(83), 4.0
    ev = T.ModelEvaluator(model=kind)
    ev.set_garbage(True)
    ev.set_move_log(True)
    ev.set_pruner(pw, act, k=k, ply_mask=mask)
    say(f"case: wide pruned k={k} mask={mask}")
    ev.stage(w, nm, seeds, cap)
    ev.launch()
    _, pl = ev.fetch()
    cnt = ev.fetch_counters()
    log = ev.fetch_moves()
    ev.close()
    full = T.ModelEvaluator(model=kind)
    full.set_garbage(True)
    full.stage(w, nm, seeds, cap)
    full.launch()
    _, full_pl = full.fetch()
    full_cnt = full.fetch_counters()
    full.close()
    # pruning cut value evaluations per placement
    assert cnt[:, 1].sum() / pl.sum() < 0.8 * full_cnt[:, 1].sum() / full_pl.sum()

    def search(e, kd, wm):
        return O.search_beam_pruned(e, kd, wm, pw, act, k, mask)

    rows = [(m, g, m * len(seeds) + g) for m in range(nm) for g in range(len(seeds))]
    with O.garbage(), ThreadPoolExecutor(16) as pool:
        same = list(pool.map(lambda t: replay_agreement(kind, w[t[0] * P:(t[0] + 1) * P], seeds[t[1]], log[t[2]],
                                                        int(pl[t[2]]), search), rows))
    rate = sum(same) / float(pl.sum())
    say(f"wide pruned placement agreement {rate:.4f} over {int(pl.sum())} moves")
    assert rate >= 0.99
