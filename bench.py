"""Benchmark: piece placements per second of one CMA-ES generation on one MI355X.

Workload (one "step" = one generation): 4 islands x 50 population x 50 games = 10,000 (model, game)
simulations, depth-4 beam search with widths 16/32/32 over every placement, games capped at 2,000
pieces, value network 16 -> 32 -> 32 -> 1 in FP32. Parameter vectors are sampled per island like a
fresh CMA-ES generation (island mean ~ N(0, 0.5^2), sigma 0.3, identity covariance). Game seeds come
from simulate()'s seed expansion. Inputs are staged into device memory before the timed region.

Prints one JSON line. `roofline` prices the dominant kernel's algorithmic integer/FP32 VALU work
(exact per-item candidate counts from the kernel x a fixed per-candidate op model) against the
MI355X VALU peak; `cpu_baseline` times the CPU oracle on a bounded sample of the same workload on
the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ISLANDS, POP, GAMES, CAP = 4, 50, 50, 2000
GEN_SEED = 0x7E7215

# Algorithmic VALU lane-ops per candidate placement of the reference algorithm (10-column board loops):
C_FIT = 120                 # spawn collision + drop distance over 10 columns
C_PLACE_CLEAR = 45          # lock-in, full-row mask, compaction
C_FEAT = {0: 355, 1: 425}   # fused column metrics + normalisation (V11 / 16-feature set)
# network: V11 = 88 multiply-adds as separate mul + add (reference rounding) + 8 ReLU;
# 16-32-32-1 = 1,633 fused multiply-adds + 64 ReLU
C_MLP = {0: 184, 1: 1697}
VALU_PEAK_TOPS = 39.3       # 256 CU x 64 lanes x 2.4 GHz (non-packed 32-bit VALU ops)


def workload(model_kind: int):
    import tetris_ai3_amd as T
    P = T.num_params(model_kind)
    rng = np.random.Generator(np.random.PCG64(GEN_SEED))
    w = np.empty((ISLANDS * POP, P), dtype=np.float32)
    for i in range(ISLANDS):
        mean = rng.normal(0.0, 0.5, P)
        w[i * POP:(i + 1) * POP] = (mean + 0.3 * rng.normal(0.0, 1.0, (POP, P))).astype(np.float32)
    seeds = T.expand_seeds(GEN_SEED, GAMES)
    return w.reshape(-1), ISLANDS * POP, seeds


def cpu_baseline(model_kind: int, w, num_models: int, seeds, budget_s: float):
    import oracle as O
    P = O.num_params(model_kind)
    threads = min(os.cpu_count() or 1, 16)
    items = threads
    t0 = time.perf_counter()
    # first `items` (model, game) items of the generation, played to completion: model-major order
    fit_all, pl_all = [], []
    done = 0
    while done < items:
        m, g = divmod(done, len(seeds))
        n = min(items - done, len(seeds) - g)
        f, pl = O.eval_host_mt(model_kind, w[m * P:(m + 1) * P], 1, seeds[g:g + n], CAP, threads=threads)
        fit_all.append(f)
        pl_all.append(pl)
        done += n
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    placements = int(np.concatenate(pl_all).sum())
    return {"value": placements / dt, "unit": "placements/s", "cores": threads, "kind": "port",
            "sample": f"first {done} (model, game) items of the generation played to completion on {threads} "
                      f"threads ({placements} placements, {dt:.1f} s)"}, np.concatenate(fit_all), done


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--model", default="f16", choices=["f16", "v11"])
    ap.add_argument("--waves-per-game", type=int, default=0, help="0: the network's tuned default")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build evaluates on a single MI355X; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import tetris_ai3_amd as T
    kind = T.MODEL_F16 if args.model == "f16" else T.MODEL_V11
    w, M, seeds = workload(kind)
    ev = T.ModelEvaluator(model=kind)
    if args.waves_per_game:
        ev.set_launch(args.waves_per_game, True, 0)
    ev.stage(w, M, seeds, CAP)

    for _ in range(args.warmup):
        ev.launch()
        ev.sync()
    kernel_ms = []
    ev.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev.launch()
        ev.sync()
        kernel_ms.append(ev.last_kernel_ms())
    elapsed = time.perf_counter() - t0
    fit, pl = ev.fetch()
    cnt = ev.fetch_counters()
    info = ev.launch_info()

    placements = int(pl.sum())
    ms_per_step = elapsed * 1000.0 / args.steps
    value = placements * args.steps / elapsed
    applies, legal = int(cnt[:, 0].sum()), int(cnt[:, 1].sum())
    ops = applies * C_FIT + legal * (C_PLACE_CLEAR + C_FEAT[kind] + C_MLP[kind])
    kms = float(np.mean(kernel_ms))
    achieved = ops / (kms / 1000.0) / 1e12
    lengths = np.sort(pl)

    out = {
        "metric": "piece-placements/sec per gen on 1 MI355X, pruned vs full beam at <=1% score loss",
        "value": value,
        "unit": "placements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic (CMA-ES-sampled random networks, seeded 7-bag games)",
        "config": {
            "workload": "one CMA-ES generation: 4 islands x 50 pop x 50 games = 10,000 simulations, depth-4 "
                        "full beam 16/32/32, cap 2,000 pieces, FP32 network "
                        + ("16-32-32-1" if kind == T.MODEL_F16 else "V11 10-8-1"),
            "beam": "full",
            "items": int(pl.size),
            "placements_per_step": placements,
            "games_per_s": pl.size * args.steps / elapsed,
            "game_length_p50": float(np.percentile(lengths, 50)),
            "game_length_p99": float(np.percentile(lengths, 99)),
            "game_length_max": int(lengths[-1]),
            "mean_score": float(fit.mean()),
            "candidates_per_placement": applies / max(placements, 1),
            "launch": info,
            "kernel_ms": kms,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": VALU_PEAK_TOPS,
            "unit": "TOP/s",
            "frac": achieved / VALU_PEAK_TOPS,
            "traffic": None,
            "ops_per_launch": ops,
        },
    }
    if not args.no_cpu:
        base, cpu_fit, n = cpu_baseline(kind, w, M, seeds, args.cpu_budget)
        out["cpu_baseline"] = base
        out["cpu_parity_items"] = n
        out["cpu_parity_ok"] = bool(cpu_fit.tobytes() == fit[:n].tobytes())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
