"""CPU: the agreement report's comparison logic, on device data made from the oracle's own games with one
planted difference (the report must count exactly the moves before it and no more)."""
import importlib
import os
import sys

import numpy as np

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def _report(monkeypatch):
    import bench
    monkeypatch.setattr(bench, "POP", 2)
    monkeypatch.setattr(bench, "GAMES", 3)
    monkeypatch.setattr(bench, "CAP", 60)
    monkeypatch.setattr(bench, "WIDE_MODELS", 2)
    monkeypatch.setattr(bench, "WIDE_GAMES", 2)
    monkeypatch.setattr(bench, "WIDE_CAP", 10)
    import agreement_report
    return importlib.reload(agreement_report), bench


def _fake_device(A, kind, w, M, seeds, cap):
    P = O.num_params(kind)
    items = M * len(seeds)
    fit = np.zeros(items, np.float32)
    pl = np.zeros(items, np.uint32)
    log = np.zeros((items, cap), np.uint8)
    for i in range(items):
        m, g = divmod(i, len(seeds))
        mv, _, f = O.play_trace(kind, w[m * P:(m + 1) * P], int(seeds[g]), cap)
        e = A.encode(mv)
        fit[i], pl[i], log[i, :e.size] = f, e.size, e
    return fit, pl, log


def test_exact_comparison_counts_moves_up_to_the_first_difference(monkeypatch, tmp_path):
    A, bench = _report(monkeypatch)
    w, M, seeds = bench.workload(1)
    fit, pl, log = _fake_device(A, 1, w, M, seeds, bench.CAP)
    victim = int(np.argmax(pl))
    assert pl[victim] > 5
    log[victim, 4] ^= 1  # a different move at index 4
    path = tmp_path / "fake.npz"
    np.savez_compressed(path, model="f16", fitness=fit, placements=pl, items=np.arange(pl.size), moves=log,
                        weights_sum=np.float64(np.asarray(w, dtype=np.float64).sum()))
    r = A.compare_exact(np.load(path), "f16", 2)
    assert r["games"] == pl.size and r["games_agreeing"] == pl.size - 1
    assert r["placements_agreeing"] == int(pl.sum()) - (int(pl[victim]) - 4)
    assert r["score_device"] == r["score_oracle"]


def test_wide_replay_agrees_with_the_oracles_own_games(monkeypatch, tmp_path):
    A, bench = _report(monkeypatch)
    w, M, seeds = bench.wide_workload()
    with O.garbage():
        fit, pl, log = _fake_device(A, 3, w, M, seeds, bench.WIDE_CAP)
    items = np.array([0, M * len(seeds) - 1])
    path = tmp_path / "fakew.npz"
    np.savez_compressed(path, model="wide", fitness=fit, placements=pl, items=items, moves=log[items, :6],
                        weights_sum=np.float64(np.asarray(w, dtype=np.float64).sum()))
    r = A.compare_wide(np.load(path), 2)
    assert r["moves_compared"] == int(np.minimum(pl[items], 6).sum())
    assert r["moves_agreeing"] == r["moves_compared"]
