"""Benchmark: piece placements per second of one CMA-ES generation on one MI355X.

Workload (one "step" = one generation): 4 islands x 50 population x 50 games = 10,000 (model, game)
simulations, depth-4 beam search with widths 16/32/32 over every placement, games capped at 2,000
pieces, value network 16 -> 32 -> 32 -> 1 in FP32. Parameter vectors are sampled per island like a
fresh CMA-ES generation (island mean ~ N(0, 0.5^2), sigma 0.3, identity covariance). Game seeds come
from simulate()'s seed expansion. Inputs are staged into device memory before the timed region; the launch
shape is autotuned on the batch first (untimed; results never depend on it).

Prints one JSON line. `value` is the full beam's rate over the K timed generations. `pruned` reports the learned
pruner: trained (untimed) on full-beam traces of this generation, then judged on HELD-OUT game seeds -- the same
population on a new seed set -- against the full beam there, at each setting of PRUNE_SPECS; beside it the
non-learned way to spend fewer value evaluations, narrower beams, and a narrower beam with the pruner. `pruned`
holds the fastest learned setting within the <=1% mean-score bar (else the one with the smallest loss),
`pruned.narrow_best` the same for narrower beams, `pruned.champion` the same study on a trained 16-32-32-1
champion (its own pruner from its own games on training seeds, judged on 1,000 held-out games). `roofline` prices
the dominant kernel's algorithmic integer/FP32 VALU work (exact per-item counts of the candidates the kernel fits
and scores x a fixed per-candidate op model) against the MI355X VALU peak; for 16-32-32-1 the hidden layers'
multiply-adds run on FP32 MFMA and are priced there (`mfma_fp32`), and the bound is the pipe with the longer
ideal time. The kernel scores each repeated list
once (a list whose parent board, clear histogram, piece and held-is-I equal an earlier list's at the same level:
the reference scores it again) and takes lists the previous move scored (its roots, and for 16-32-32-1 lists of
levels 1 and 2) from that move's scores, so these counts are the reference's less its repeats and those lists;
the CPU leg reports both shares on its sample. `traffic` is the profiled HBM bytes per launch when a PMC summary for the workload is
committed under profiles/. `cpu_baseline` times the CPU oracle on a fixed sample of the same workload (items spread
over the generation) on all host cores, and a smaller sample on one thread; the sampled items' fitness,
placements and counters must equal the GPU's.

`--workload wide` measures the scaled configuration instead: 256 models x 100 games, cap 5,000 pieces,
32-128-128-1 network on BF16 MFMA, garbage rows, with its own pruned leg (pruner trained on that generation's
traces, judged on held-out seeds).
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

# Algorithmic VALU lane-ops per candidate placement, frozen from the instrumented scalar restatement of one
# candidate (oracle/opcount.cpp; tests/test_oracle_opcount.py re-derives them): every scalar operation of fit
# (over the columns the piece occupies), place + clear, the fused feature pass and the dense network, averaged over
# the legal candidates of full-beam games (fit and place + clear depend on the data, the rest does not).
C_FIT = 54                   # spawn collision + drop distance, per fitted candidate
C_PLACE_CLEAR = {0: 33, 1: 32}  # lock-in, full-row mask, row compaction, histogram bump (V11 / 16-32-32-1 games)
C_FEAT = {0: 470, 1: 611}    # fused column metrics + normalisation (V11 / 16-feature set)
# network: V11 = 88 multiply-adds as separate mul + add (reference rounding) + 8 ReLU = 184;
# 16-32-32-1 = 1,568 fused multiply-adds + 64 ReLU = 1,632 (every multiply-add of the network, including those the
# kernel skips exactly when an input is zero in every lane of a chunk)
C_MLP = {0: 184, 1: 1632}
VALU_PEAK_TOPS = 39.3       # 256 CU x 64 lanes x 2.4 GHz (non-packed 32-bit VALU ops)
# the plain 16-32-32-1 kernel runs the two hidden layers (16 x 32 + 32 x 32 multiply-adds of C_MLP) on
# v_mfma_f32_16x16x4_f32 for finite weights (every model bench.py samples); ReLU and the output layer stay on the VALU
F16_MFMA_MACS = 16 * 32 + 32 * 32
FP32_MFMA_PEAK_TFLOPS = 157.3  # dense FP32 MFMA
BF16_PEAK_TFLOPS = 2500.0   # dense BF16 MFMA
F32W_FLOPS = 2 * 20865      # per leaf evaluation of the wide network (on MFMA)
PRUNE_EPOCHS = 16
PRUNE_RECORDS = 8_000_000  # training sample of the generation's ~17 M trace records
HELDOUT_SEED = GEN_SEED + 1  # the pruner is judged on games of seeds it never saw
# settings of the pruned leg: p<k>:<mask> = learned pruner keeps k of every list at the plies of mask (bits 1, 2 =
# middle levels, 3 = leaves); n<k0>-<k1>-<k2> = narrower beam (no pruner); both joined by +
PRUNE_SPECS = ("p16:8", "p8:8", "p8:12", "p4:12", "n12-24-24", "n8-16-16", "n4-8-8", "n8-16-16+p8:8")
CHAMP_PATH = os.path.join(ROOT, "tests", "golden", "f16_champion.champ")  # a trained 16-32-32-1 model
CHAMP_TRAIN_SEED, CHAMP_HELDOUT_SEED, CHAMP_GAMES = 0x5EED, 0x5EEE, 1000
CPU_SAMPLE = 128  # items of the generation the CPU baseline plays, evenly spaced through it (model-major)


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


def _best(rows):
    """the fastest row within the 1% bar, else the one with the smallest loss"""
    ok = [r for r in rows if r["meets_1pct"]]
    return max(ok, key=lambda r: r["placements_per_s"]) if ok else min(rows, key=lambda r: r["mean_score_loss_pct"])


def _summary(rows, note):
    full, rest = rows[0], rows[1:]
    learned = [r for r in rest if r["k"] and r["widths"] == [16, 32, 32]]
    narrow = [r for r in rest if not r["k"]]
    out = dict(_best(learned)) if learned else {}
    out.update({"judged_on": note, "full_beam": full, "settings": rest})
    if narrow:
        out["narrow_best"] = _best(narrow)
    return out


def pruned_leg(T, kind, w, M, seeds, cap, steps, every=512, cap_records=24_000_000, heldout_seed=HELDOUT_SEED):
    """Pruner trained on this generation's full-beam traces (untimed); every setting judged against the full beam
    on the same population playing held-out seeds."""
    from tetris_ai3_amd import learn
    ev = T.ModelEvaluator(model=kind)
    if kind == T.MODEL_F32W:
        ev.set_garbage(True)
    pw, act, info = learn.fit_pruner(ev, w, M, seeds, cap, every=every, cap_records=cap_records,
                                     records=PRUNE_RECORDS, epochs=PRUNE_EPOCHS)
    held = T.expand_seeds(heldout_seed, len(seeds))
    rows = learn.compare_settings(ev, w, M, held, cap, PRUNE_SPECS, pw, act, steps)
    ev.close()
    out = _summary(rows, f"the same {M} models on {len(seeds)} held-out seeds (simulate seed {heldout_seed:#x})")
    out["trained_on"] = (f"{info['trained_on']} of {info['trace_records']} full-beam trace records of this "
                         f"generation (1 move in {every} per game), {info['epochs']} epochs")
    out["training_loss"] = info["training_loss"]
    out["setup_s"] = info["setup_s"]
    return out


def champion_leg(T, steps):
    """The same study on a trained 16-32-32-1 champion: its pruner trained on its own full-beam games on training
    seeds, every setting judged on 1,000 held-out games."""
    if not os.path.exists(CHAMP_PATH):
        return None
    from tetris_ai3_amd import learn
    w = T.champ_read(CHAMP_PATH).astype(np.float32)
    ev = T.ModelEvaluator(model=T.MODEL_F16)
    # the trace buffer holds every record the 200 games produce (~22.5 M), so the training set, and with it every
    # figure, does not depend on the order the device plays the games in
    pw, act, info = learn.fit_pruner(ev, w, 1, T.expand_seeds(CHAMP_TRAIN_SEED, 200), CAP, every=64,
                                     cap_records=24_000_000, records=4_000_000, epochs=10)
    rows = learn.compare_settings(ev, w, 1, T.expand_seeds(CHAMP_HELDOUT_SEED, CHAMP_GAMES), CAP, PRUNE_SPECS, pw,
                                  act, steps)
    ev.close()
    out = _summary(rows, f"{CHAMP_GAMES} held-out games (simulate seed {CHAMP_HELDOUT_SEED:#x}) of "
                         f"{os.path.relpath(CHAMP_PATH, ROOT)}")
    out["trained_on"] = (f"{info['trained_on']} of {info['trace_records']} trace records of 200 training games "
                         f"(1 move in 64), {info['epochs']} epochs")
    return out


def cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, cores


def cpu_baseline(model_kind: int, w, num_models: int, seeds, cap, gpu_fit, gpu_pl, gpu_cnt):
    """The oracle (this build's scalar C++ restatement of eval_host, model_eval_host.cpp:26-54) on CPU_SAMPLE items
    spread evenly over the generation, one item per task on every host core, each played to completion; then
    the first of those items (about 2,500 placements by the GPU's count) on one thread, as eval_host runs. The
    sampled items' fitness must equal the GPU's bit for bit, and their work counters the GPU's plus the repeated
    lists the GPU scores once."""
    import oracle as O
    from concurrent.futures import ThreadPoolExecutor
    P, G = O.num_params(model_kind), len(seeds)
    items = num_models * G
    sample = [(k * items) // CPU_SAMPLE for k in range(CPU_SAMPLE)]

    def one(i):
        m, g = divmod(i, G)
        c = O.Counters()
        f, p = O.eval_host_mt(model_kind, w[m * P:(m + 1) * P], 1, seeds[g:g + 1], cap, threads=1, counters=c)
        return f[0], int(p[0]), c

    model, cores = cpu_info()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as pool:
        res = list(pool.map(one, sample))
    dt = time.perf_counter() - t0
    placements = sum(p for _, p, _ in res)
    single, acc = [], 0
    for i in sample:
        if single and acc + int(gpu_pl[i]) > 2500:
            break
        single.append(i)
        acc += int(gpu_pl[i])
    t1 = time.perf_counter()
    res1 = [one(i) for i in single]
    dt1 = time.perf_counter() - t1
    p1 = sum(p for _, p, _ in res1)
    fit = np.array([f for f, _, _ in res], dtype=np.float32)
    ok = fit.tobytes() == gpu_fit[sample].tobytes() and [p for _, p, _ in res] == [int(gpu_pl[i]) for i in sample]
    ok = ok and all(int(gpu_cnt[i, 0]) == c.applies - c.repeat_applies - c.reuse_applies
                    and int(gpu_cnt[i, 1]) == c.legal - c.repeat_legal - c.reuse_legal
                    for i, (_, _, c) in zip(sample, res))
    legal = sum(c.legal for _, _, c in res)
    repeats = {"scored_by_the_reference_search": legal,
               "in_repeated_lists": sum(c.repeat_legal for _, _, c in res),
               "share": sum(c.repeat_legal for _, _, c in res) / max(legal, 1),
               "taken_from_the_previous_move": sum(c.reuse_legal for _, _, c in res),
               "share_taken": sum(c.reuse_legal for _, _, c in res) / max(legal, 1),
               "sample": f"the {CPU_SAMPLE} sampled items, counted by the oracle"}
    return {"value": placements / dt, "unit": "placements/s", "cores": cores, "kind": "port",
            "sample": f"{CPU_SAMPLE} (model, game) items evenly spaced over the generation (item k * {items} / "
                      f"{CPU_SAMPLE}), each played to completion, one per task on {cores} threads "
                      f"({placements} placements, {dt:.1f} s)",
            "one_thread": {"value": p1 / dt1, "unit": "placements/s",
                           "sample": f"the first {len(single)} of those items on 1 thread ({p1} placements, {dt1:.1f} s)"},
            "nproc": os.cpu_count(), "cpu_model": model,
            "build": "g++ -O3 -march=x86-64-v3 -ffp-contract=off (oracle/Makefile)"}, len(sample), ok, repeats


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
    ap.add_argument("--waves-per-game", type=int, default=0,
                    help="fixed waves per game (1, 2, 4) instead of the autotuned launch shape")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prune-steps", type=int, default=3, help="timed launches per pruned-leg setting")
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
    tuned = None
    if args.waves_per_game:
        ev.set_launch(args.waves_per_game, True, 0)
        ev.stage(w, M, seeds, cap)
    else:  # launch shape and first-phase budget tuned on this batch (untimed; results never depend on them)
        tuned = T.autotune(ev, w, M, seeds, cap, candidates=((2, 0), (4, 0)), phases=(None, 32), use_cache=False)

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
    else:  # VALU lane-ops; for 16-32-32-1 the hidden layers' multiply-adds are priced on the FP32 MFMA peak
        macs = legal * F16_MFMA_MACS if kind == T.MODEL_F16 else 0
        ops = applies * C_FIT + legal * (C_PLACE_CLEAR[kind] + C_FEAT[kind] + C_MLP[kind]) - macs
        achieved = ops / (kms / 1000.0) / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                "frac": achieved / VALU_PEAK_TOPS,
                "traffic": traffic_from_profiles(f"traffic_{'f16' if kind == T.MODEL_F16 else 'v11'}_generation"),
                "ops_per_launch": ops}
        if macs:  # the two pipes run side by side (MFMA of one wave, VALU of another): the slower ideal is the bound
            mf = 2 * macs / (kms / 1000.0) / 1e12
            mfma = {"achieved": mf, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": mf / FP32_MFMA_PEAK_TFLOPS, "flops_per_launch": 2 * macs}
            if mfma["frac"] > roof["frac"]:
                roof = {"bound": "mfma", **{k: mfma[k] for k in ("achieved", "peak", "unit", "frac")},
                        "traffic": roof["traffic"], "flops_per_launch": 2 * macs,
                        "valu": {k: roof[k] for k in ("achieved", "peak", "unit", "frac", "ops_per_launch")}}
            else:
                roof["mfma_fp32"] = mfma
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
            "autotune": tuned,
            "kernel_ms": kms,
        },
        "roofline": roof,
    }
    if not args.no_pruned:
        if wide:
            out["pruned"] = pruned_leg(T, kind, w, M, seeds, cap, args.prune_steps, every=1024,
                                       heldout_seed=GEN_SEED + 5)
        else:
            out["pruned"] = pruned_leg(T, kind, w, M, seeds, cap, args.prune_steps)
            champ = champion_leg(T, args.prune_steps) if kind == T.MODEL_F16 else None
            if champ is not None:
                out["pruned"]["champion"] = champ
    if not args.no_cpu and not wide:
        base, n, ok, repeats = cpu_baseline(kind, w, M, seeds, cap, fit, pl, cnt)
        out["cpu_baseline"] = base
        out["cpu_parity_items"] = n
        out["cpu_parity_ok"] = bool(ok)
        out["config"]["value_evaluations_in_repeated_lists"] = repeats
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
