"""Move-by-move agreement of the HIP evaluator with the CPU oracle, over a whole generation.

Two steps, because the oracle needs hours of CPU time where the device needs seconds:

    python tests/agreement_report.py gpu agreement_f16.npz [--model f16|v11|wide]   # with the GPU
    python tests/agreement_report.py compare agreement_f16.npz [--out report.json]  # on host cores

`gpu` plays bench.py's workload with the device's move log on and stores every item's fitness, placement count
and moves (hold, rotation, x of each placement).
- f16 / v11: the 10,000-game generation (4 islands x 50 population x 50 games, cap 2,000, depth-4 beam 16/32/32).
  `compare` plays every game again on the oracle and compares it with the device's move by move. A game agrees
  when its placement count, every move and its fitness bits are the oracle's; the placement-agreement rate counts
  the moves that equal the oracle's up to each game's first difference, over all moves.
- wide: the scaled workload (256 models x 100 games, cap 5,000, garbage rows, 32-128-128-1 on BF16 MFMA against
  the oracle's FP32 restatement), on an evenly spread sample of games and their first moves (the oracle takes
  ~0.1 s per move there). Its games may part from the oracle's (the two networks round differently), so the
  sample is replayed along the device's own moves and, at every state the device reached, the oracle's search
  picks its own move; the rate counts the states where both pick the same placement.
Both report the score distribution of the device's games (and the oracle's, for the exact networks).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402  (the workload definitions)

KINDS = {"v11": 0, "f16": 1, "wide": 3}


def _workload(model: str):
    if model == "wide":
        w, M, seeds = bench.wide_workload()
        return w, M, seeds, bench.WIDE_CAP
    w, M, seeds = bench.workload(KINDS[model])
    return w, M, seeds, bench.CAP


def run_gpu(out: str, model: str, sample: int, moves: int, min_length: int) -> None:
    import tetris_ai3_amd as T
    kind = KINDS[model]
    w, M, seeds, cap = _workload(model)
    ev = T.ModelEvaluator(model=kind)
    ev.set_garbage(model == "wide")
    ev.set_move_log(True)
    print(f"gpu: {model} {M} models x {len(seeds)} games cap {cap}", flush=True)
    ev.stage(w, M, seeds, cap)
    ev.launch()
    fit, pl = ev.fetch()
    log = ev.fetch_moves()
    ev.close()
    items = np.arange(M * len(seeds))
    if model == "wide":  # an evenly spread sample of the games at least min_length long, first `moves` moves
        pool = np.nonzero(pl >= min_length)[0]
        items = pool[np.linspace(0, pool.size - 1, min(sample, pool.size)).round().astype(np.int64)]
        log = log[items, :moves]
    np.savez_compressed(out, model=model, fitness=fit, placements=pl, items=items, moves=log,
                        weights_sum=np.float64(np.asarray(w, dtype=np.float64).sum()))
    print(f"gpu: {int(pl.sum())} placements; {log.size} move-log bytes of {items.size} items written to {out}",
          flush=True)


def distribution(fit: np.ndarray) -> dict:
    f = fit.astype(np.float64)
    q = np.percentile(f, [1, 10, 25, 50, 75, 90, 99])
    return {"mean": float(f.mean()), "std": float(f.std()), "min": float(f.min()), "p1": q[0], "p10": q[1],
            "p25": q[2], "p50": q[3], "p75": q[4], "p90": q[5], "p99": q[6], "max": float(f.max())}


def encode(moves: np.ndarray) -> np.ndarray:
    """oracle game record rows (hold, rotation, x + 2, cleared) -> move-log bytes"""
    m = moves.astype(np.uint32)
    return ((m[:, 0] << 7) | (m[:, 1] << 5) | m[:, 2]).astype(np.uint8)


def _progress(done: int, total: int, t0: float, every: int) -> None:
    if done % every == 0:
        print(f"compare: {done}/{total} items, {time.perf_counter() - t0:.0f} s", flush=True)


def compare_exact(d, model: str, threads: int, every: int = 500) -> dict:
    """f16 / v11: every game replayed on the oracle, compared move by move"""
    import oracle as O
    kind = KINDS[model]
    w, M, seeds, cap = _workload(model)
    P = O.num_params(kind)
    gfit, gpl, glog = d["fitness"], d["placements"], d["moves"]
    items = M * len(seeds)
    assert gfit.shape == (items,) and glog.shape == (items, cap)
    hfit = np.zeros(items, dtype=np.float32)
    hpl = np.zeros(items, dtype=np.uint32)
    prefix = np.zeros(items, dtype=np.int64)  # moves equal up to the first difference
    same_moves = np.zeros(items, dtype=bool)

    def one(i: int) -> None:
        m, g = divmod(i, len(seeds))
        mv, _, f = O.play_trace(kind, w[m * P:(m + 1) * P], int(seeds[g]), cap)
        hm = encode(mv)
        hfit[i], hpl[i] = f, hm.size
        n = min(int(gpl[i]), hm.size)
        diff = np.nonzero(glog[i, :n] != hm[:n])[0]
        prefix[i] = diff[0] if diff.size else n
        same_moves[i] = diff.size == 0 and int(gpl[i]) == hm.size

    order = np.argsort(-gpl.astype(np.int64), kind="stable")  # longest games first
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as pool:
        for done, _ in enumerate(pool.map(one, order.tolist()), 1):
            _progress(done, items, t0, every)
    same = same_moves & (gfit.view(np.uint32) == hfit.view(np.uint32))
    return {
        "workload": f"{M} models x {len(seeds)} games, cap {cap}, depth-4 beam 16/32/32, {model}",
        "games": items,
        "games_agreeing": int(same.sum()),
        "game_agreement_rate": float(same.mean()),
        "placements_device": int(gpl.sum()),
        "placements_oracle": int(hpl.sum()),
        "placements_agreeing": int(prefix.sum()),
        "placement_agreement_rate": float(prefix.sum() / max(1, gpl.sum())),
        "score_device": distribution(gfit),
        "score_oracle": distribution(hfit),
        "game_length": {"p50": float(np.percentile(gpl, 50)), "p99": float(np.percentile(gpl, 99)),
                        "mean": float(gpl.mean()), "at_cap": int((gpl == cap).sum())},
        "oracle_threads": threads,
        "oracle_seconds": round(time.perf_counter() - t0, 1),
    }


def compare_wide(d, threads: int) -> dict:
    """wide: the sampled device games replayed on the oracle's engine, the oracle's move at every state"""
    import oracle as O
    kind = KINDS["wide"]
    w, M, seeds, cap = _workload("wide")
    P = O.num_params(kind)
    gfit, gpl, items, glog = d["fitness"], d["placements"], d["items"], d["moves"]
    n_moves = np.minimum(gpl[items].astype(np.int64), glog.shape[1])
    same = np.zeros(items.size, dtype=np.int64)

    def one(j: int) -> None:
        i = int(items[j])
        m, g = divmod(i, len(seeds))
        e = O.engine(int(seeds[g]))
        for t in range(int(n_moves[j])):
            b = int(glog[j, t])
            hold, o, x = b >> 7, (b >> 5) & 3, (b & 15) - 2
            r = O.search_beam(e, kind, w[m * P:(m + 1) * P])
            if r is not None and (int(r["hold"]), r["o"], r["x"]) == (hold, o, x):
                same[j] += 1
            if hold:
                O.engine_hold(e)
            if O.engine_place(e, o, x) < 0:
                break

    t0 = time.perf_counter()
    with O.garbage(), ThreadPoolExecutor(threads) as pool:
        for done, _ in enumerate(pool.map(one, range(items.size)), 1):
            _progress(done, items.size, t0, 8)
    return {
        "workload": f"{M} models x {len(seeds)} games, cap {cap}, garbage rows, 32-128-128-1 on BF16 MFMA; "
                    f"{items.size} games (at least {int(gpl[items].min())} moves long) evenly spread over the batch, "
                    f"their first {glog.shape[1]} moves",
        "games_sampled": int(items.size),
        "moves_compared": int(n_moves.sum()),
        "moves_agreeing": int(same.sum()),
        "placement_agreement_rate": float(same.sum() / max(1, n_moves.sum())),
        "sampled_games_fully_agreeing": int((same == n_moves).sum()),
        "score_device": distribution(gfit),
        "game_length": {"p50": float(np.percentile(gpl, 50)), "p99": float(np.percentile(gpl, 99)),
                        "mean": float(gpl.mean()), "at_cap": int((gpl == cap).sum())},
        "oracle_threads": threads,
        "oracle_seconds": round(time.perf_counter() - t0, 1),
    }


def run_compare(path: str, threads: int, out: str | None, every: int) -> int:
    d = np.load(path)
    model = str(d["model"])
    w, _, _, _ = _workload(model)
    assert float(np.asarray(w, dtype=np.float64).sum()) == float(d["weights_sum"]), "workload differs"
    report = compare_wide(d, threads) if model == "wide" else compare_exact(d, model, threads, every)
    text = json.dumps(report, indent=1)
    print(text, flush=True)
    if out:
        with open(out, "w") as f:
            f.write(text + "\n")
    if model == "wide":
        return 0
    return 0 if report["games_agreeing"] == report["games"] else 1


def main() -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("gpu")
    g.add_argument("out")
    g.add_argument("--model", default="f16", choices=sorted(KINDS))
    g.add_argument("--sample", type=int, default=64, help="wide: games in the sample")
    g.add_argument("--moves", type=int, default=150, help="wide: moves per sampled game")
    g.add_argument("--min-length", type=int, default=0, help="wide: sample only games at least this long")
    c = sub.add_parser("compare")
    c.add_argument("path")
    c.add_argument("--threads", type=int, default=min(os.cpu_count() or 1, 16))
    c.add_argument("--out", default=None, help="also write the JSON report here")
    c.add_argument("--progress", type=int, default=500, help="print a progress line every this many games")
    a = ap.parse_args()
    if a.cmd == "gpu":
        run_gpu(a.out, a.model, a.sample, a.moves, a.min_length)
        return 0
    return run_compare(a.path, a.threads, a.out, a.progress)


if __name__ == "__main__":
    sys.exit(main())
