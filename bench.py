"""Benchmark: piece placements per second of one CMA-ES generation on one MI355X.

Workload (one "step" = one generation): 4 islands x 50 population x 50 games = 10,000 (model, game)
simulations, depth-4 beam search with widths 16/32/32 over every placement, games capped at 2,000
pieces, value network 16 -> 32 -> 32 -> 1 in FP32. Parameter vectors are sampled per island like a
fresh CMA-ES generation (island mean ~ N(0, 0.5^2), sigma 0.3, identity covariance). Game seeds come
from simulate()'s seed expansion. Inputs are staged into device memory before the timed region.

Prints one JSON line. `value` is the full beam's rate over the K timed generations. `pruned` reports the same
generation searched with the learned pruner: the pruner is trained first (untimed) on full-beam traces of this
generation, then the pruned generation is timed at each (k, plies) setting of PRUNE_SETTINGS and its mean
score compared with the full beam's (the <=1% score-loss bar); `pruned` holds the fastest setting within the
bar, or the one with the smallest loss, and `pruned.settings` all of them. `roofline` prices the dominant kernel's
algorithmic integer/FP32 VALU work (exact per-item candidate counts from the kernel x a fixed per-candidate op
model) against the MI355X VALU peak; `traffic` is the profiled HBM bytes per launch when a PMC summary for the
workload is committed under profiles/. `cpu_baseline` times the CPU oracle on a bounded sample of the same
workload on the host cores.

`--workload wide` measures the scaled configuration instead: 256 models x 100 games, cap 5,000 pieces,
32-128-128-1 network on BF16 MFMA, garbage rows.
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
WIDE_MODELS, WIDE_GAMES, WIDE_CAP = 256, 100, 5000

# Algorithmic VALU lane-ops per candidate placement of the reference algorithm (10-column board loops):
C_FIT = 120                 # spawn collision + drop distance over 10 columns
C_PLACE_CLEAR = 45          # lock-in, full-row mask, compaction
C_FEAT = {0: 355, 1: 425}   # fused column metrics + normalisation (V11 / 16-feature set)
# network: V11 = 88 multiply-adds as separate mul + add (reference rounding) + 8 ReLU;
# 16-32-32-1 = 1,633 fused multiply-adds + 64 ReLU
C_MLP = {0: 184, 1: 1697}
VALU_PEAK_TOPS = 39.3       # 256 CU x 64 lanes x 2.4 GHz (non-packed 32-bit VALU ops)
BF16_PEAK_TFLOPS = 2500.0   # dense BF16 MFMA
F32W_FLOPS = 2 * 20865      # per leaf evaluation of the wide network (on MFMA)
PRUNE_EPOCHS = 16
PRUNE_RECORDS = 8_000_000  # training sample of the generation's ~17 M trace records
# (k, ply mask) settings of the pruned leg: mask bits 1, 2 = middle levels, 3 = leaves
PRUNE_SETTINGS = ((8, 0b1100), (8, 0b1000), (16, 0b1000), (16, 0b1100))


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


def wide_workload():
    import tetris_ai3_amd as T
    P = T.num_params(T.MODEL_F32W)
    rng = np.random.Generator(np.random.PCG64(GEN_SEED + 4))
    w = (rng.normal(0.0, 0.5, P) + 0.3 * rng.normal(0.0, 1.0, (WIDE_MODELS, P))).astype(np.float32)
    return w.reshape(-1), WIDE_MODELS, T.expand_seeds(GEN_SEED + 4, WIDE_GAMES)


def traffic_from_profiles(name: str):
    """HBM bytes per launch from a committed PMC summary (profiles/<name>.json), or None"""
    path = os.path.join(ROOT, "profiles", f"{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def pruned_leg(T, kind, w, M, seeds, full_fit, full_pl, full_ms, steps):
    """Trains the pruner on this generation's full-beam traces (untimed), then times the pruned generation at
    each (k, ply mask) of PRUNE_SETTINGS. Reports every setting and picks the fastest one within 1% mean-score
    loss of the full beam, or the one with the smallest loss when none is."""
    from tetris_ai3_amd import learn
    ev = T.ModelEvaluator(model=kind)
    t0 = time.perf_counter()
    # a buffer for every record (~17 M of 44 B): the kept set, the sample and the trained pruner are reproducible
    recs, produced, _ = learn.collect_traces(ev, w, M, seeds, CAP, every=512, cap=24_000_000)
    rng = np.random.default_rng(1)
    train = recs if recs.shape[0] <= PRUNE_RECORDS else recs[rng.choice(recs.shape[0], PRUNE_RECORDS, replace=False)]
    pw, losses, _ = learn.train_pruner(ev, train, epochs=PRUNE_EPOCHS, batch=512, lr=1e-3, seed=1)
    act = learn.calibrate_act_range(ev, pw, train)
    setup_s = time.perf_counter() - t0
    full_rate = float(full_pl.sum()) / (full_ms / 1000.0)
    settings = []
    for k, mask in PRUNE_SETTINGS:
        ev.set_pruner(pw, act, k=k, ply_mask=mask)
        ev.stage(w, M, seeds, CAP)
        ev.launch()
        ev.sync()
        t1 = time.perf_counter()
        for _ in range(steps):
            ev.launch()
            ev.sync()
        elapsed = time.perf_counter() - t1
        fit, pl = ev.fetch()
        cnt = ev.fetch_counters()
        rate = float(pl.sum()) * steps / elapsed
        loss = 1.0 - float(fit.mean()) / float(full_fit.mean()) if full_fit.mean() > 0 else \
            float(full_fit.mean() - fit.mean()) / abs(float(full_fit.mean()))
        settings.append({"k": k, "ply_mask": mask, "placements_per_s": rate, "ms_per_step": elapsed * 1000 / steps,
                         "speedup_vs_full": rate / full_rate, "mean_score": float(fit.mean()),
                         "mean_score_loss_pct": 100.0 * loss, "meets_1pct": bool(loss <= 0.01),
                         "value_evals_per_placement": float(cnt[:, 1].sum()) / max(float(pl.sum()), 1.0)})
    ev.close()
    ok = [x for x in settings if x["meets_1pct"]]
    best = max(ok, key=lambda x: x["placements_per_s"]) if ok else min(settings, key=lambda x: x["mean_score_loss_pct"])
    out = dict(best)
    out.update({"settings": settings,
                "trained_on": f"{train.shape[0]} of {produced} full-beam trace records of this generation "
                              f"(1 move in 512 per game{'' if recs.shape[0] == produced else '; buffer full'}), "
                              f"{PRUNE_EPOCHS} epochs",
                "training_loss": [float(x) for x in losses],
                "setup_s": setup_s})
    return out


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
    ap.add_argument("--workload", default="generation", choices=["generation", "wide"])
    ap.add_argument("--no-pruned", action="store_true", help="skip the pruned-beam leg")
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="copy 1 GiB device-to-device first: a kernel of known traffic to scale FETCH_SIZE / "
                         "WRITE_SIZE by in the same profiled run")
    ap.add_argument("--waves-per-game", type=int, default=0, help="0: the network's tuned default")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build evaluates on a single MI355X; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.pmc_calibration:
        import torch
        src = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
    import tetris_ai3_amd as T
    wide = args.workload == "wide"
    if wide:
        kind = T.MODEL_F32W
        w, M, seeds = wide_workload()
        cap = WIDE_CAP
    else:
        kind = T.MODEL_F16 if args.model == "f16" else T.MODEL_V11
        w, M, seeds = workload(kind)
        cap = CAP
    ev = T.ModelEvaluator(model=kind)
    if wide:
        ev.set_garbage(True)
    if args.waves_per_game:
        ev.set_launch(args.waves_per_game, True, 0)
    ev.stage(w, M, seeds, cap)

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
    applies, legal, lanes = int(cnt[:, 0].sum()), int(cnt[:, 1].sum()), int(cnt[:, 2].sum())
    kms = float(np.mean(kernel_ms))
    lengths = np.sort(pl)
    if wide:  # the wide network runs on MFMA: price its flops against the BF16 peak
        ops = legal * F32W_FLOPS
        achieved = ops / (kms / 1000.0) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / BF16_PEAK_TFLOPS, "traffic": traffic_from_profiles("traffic_f32w_wide"),
                "flops_per_launch": ops}
    else:
        ops = applies * C_FIT + legal * (C_PLACE_CLEAR + C_FEAT[kind] + C_MLP[kind])
        achieved = ops / (kms / 1000.0) / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                "frac": achieved / VALU_PEAK_TOPS,
                "traffic": traffic_from_profiles(f"traffic_{'f16' if kind == T.MODEL_F16 else 'v11'}_generation"),
                "ops_per_launch": ops}
    if wide:
        workload_name = ("scaled generation: 256 models x 100 games = 25,600 simulations, depth-4 full beam "
                         "16/32/32, cap 5,000 pieces, garbage rows 0-4, 32-128-128-1 network on BF16 MFMA")
    else:
        workload_name = ("one CMA-ES generation: 4 islands x 50 pop x 50 games = 10,000 simulations, depth-4 "
                         "full beam 16/32/32, cap 2,000 pieces, FP32 network "
                         + ("16-32-32-1" if kind == T.MODEL_F16 else "V11 10-8-1"))

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
        "dtype": "int32+bf16" if wide else "int32+fp32",
        "data": "synthetic (CMA-ES-sampled random networks, seeded 7-bag games)",
        "config": {
            "workload": workload_name,
            "beam": "full",
            "items": int(pl.size),
            "placements_per_step": placements,
            "games_per_s": pl.size * args.steps / elapsed,
            "game_length_p50": float(np.percentile(lengths, 50)),
            "game_length_p99": float(np.percentile(lengths, 99)),
            "game_length_max": int(lengths[-1]),
            "mean_score": float(fit.mean()),
            "candidates_per_placement": applies / max(placements, 1),
            "lane_utilisation": legal / max(lanes, 1),
            "launch": info,
            "kernel_ms": kms,
        },
        "roofline": roof,
    }
    if not wide and not args.no_pruned:
        out["pruned"] = pruned_leg(T, kind, w, M, seeds, fit, pl, ms_per_step, args.steps)
    if not args.no_cpu and not wide:
        base, cpu_fit, n = cpu_baseline(kind, w, M, seeds, args.cpu_budget)
        out["cpu_baseline"] = base
        out["cpu_parity_items"] = n
        out["cpu_parity_ok"] = bool(cpu_fit.tobytes() == fit[:n].tobytes())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
