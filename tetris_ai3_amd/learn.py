"""Learned pruner pipeline: full-beam traces -> training on the GPU -> calibration -> pruned beam.

Every step runs in libta3h.so on the GPU (trace emission inside the search kernel, forward/backward/Adam
kernels, MFMA inference inside the pruned search); this module only sequences the calls.

Labels (emitted by the full beam itself): a ply-1 or ply-2 candidate is positive when the level's top-K kept
it, a ply-3 candidate when it is its list's best leaf. The pruner ranks the legal placements of every
(kept node, piece) list; the pruned search scores only the top k of them with the value network.
"""
from __future__ import annotations

import time

import numpy as np

from . import ModelEvaluator, pruner_init

PLY_SHIFT, LABEL_SHIFT = 9, 11


def canonical_order(recs):
    """The records sorted by content. The device appends them in no particular order and training batches follow
    the record order, so sorting makes training a function of the record set alone. Key: a 64-bit hash of the 11
    words (records with equal keys are equal, but for a hash collision, odds ~n^2 / 2^65)."""
    recs = np.ascontiguousarray(recs, dtype=np.uint32)
    h = np.zeros(recs.shape[0], dtype=np.uint64)
    for j in range(recs.shape[1]):  # multiply-xorshift mixing, wrapping mod 2^64
        h = (h ^ recs[:, j].astype(np.uint64)) * np.uint64(0x9E3779B97F4A7C15)
        h ^= h >> np.uint64(31)
    return recs[np.argsort(h, kind="stable")]


def collect_traces(ev: ModelEvaluator, weights, num_models: int, seeds, max_moves: int, every: int, cap: int):
    """Plays the batch with the full beam and returns (records uint32 [n, 11] in canonical order, produced,
    fitness). Which records a full buffer keeps depends on the order of work: with cap >= produced the set, and
    so everything trained from it, is reproducible."""
    ev.set_trace(cap, every)
    ev.stage(weights, num_models, seeds, max_moves)
    ev.launch()
    fit, _ = ev.fetch()
    recs, produced = ev.fetch_trace()
    ev.set_trace(0)
    return canonical_order(recs), produced, fit


def label_stats(recs) -> dict:
    ply = (recs[:, 10] >> PLY_SHIFT) & 3
    lab = (recs[:, 10] >> LABEL_SHIFT) & 1
    return {int(p): (int((ply == p).sum()), float(lab[ply == p].mean()) if (ply == p).any() else 0.0)
            for p in (1, 2, 3)}


def calibrate_act_range(ev: ModelEvaluator, w, recs, quantile: float = 0.9999, max_samples: int = 65536,
                        seed: int = 0) -> float:
    """Upper range of the first hidden layer for the 8-bit activation grid: the given quantile of the
    positive activations over a sample of the records."""
    rng = np.random.default_rng(seed)
    smp = recs if recs.shape[0] <= max_samples else recs[rng.choice(recs.shape[0], max_samples, replace=False)]
    h = ev.pruner_hidden(w, smp)
    pos = h[h > 0]
    return float(np.quantile(pos, quantile)) if pos.size else 1.0


def train_pruner(ev: ModelEvaluator, recs, epochs: int = 3, batch: int = 512, lr: float = 1e-3,
                 pos_weight: float | None = None, seed: int = 0, init=None):
    """Adam on the device. pos_weight defaults to the negative/positive ratio of the records.
    Returns (weights, per-epoch losses, seconds)."""
    lab = (recs[:, 10] >> LABEL_SHIFT) & 1
    if pos_weight is None:
        npos = max(int(lab.sum()), 1)
        pos_weight = float(lab.size - npos) / npos
    w0 = pruner_init(seed) if init is None else np.asarray(init, dtype=np.float32)
    t0 = time.perf_counter()
    w, losses = ev.pruner_train(w0, recs, epochs=epochs, batch=batch, lr=lr, pos_weight=pos_weight, seed=seed)
    return w, losses, time.perf_counter() - t0


FULL_WIDTHS = (16, 32, 32)


def parse_setting(spec: str) -> dict:
    """'p<k>:<mask>' = pruned beam (k of every list at the plies of mask: bit 1, 2 middle levels, bit 3 leaves),
    'n<k0>-<k1>-<k2>' = narrower beam, 'n...+p...' both, 'full' = the full beam"""
    out = {"name": spec, "widths": FULL_WIDTHS, "k": 0, "mask": 0}
    for part in spec.split("+"):
        if part == "full":
            continue
        if part.startswith("p"):
            k, mask = part[1:].split(":")
            out["k"], out["mask"] = int(k), int(mask)
        elif part.startswith("n"):
            out["widths"] = tuple(int(v) for v in part[1:].split("-"))
        else:
            raise ValueError(f"unknown setting {spec!r}")
    return out


def run_setting(ev: ModelEvaluator, w, num_models: int, seeds, max_moves: int, setting: dict, pruner=None,
                act_range: float = 0.0, steps: int = 3) -> dict:
    """One warm launch, then `steps` timed launches (wall clock, launch + synchronisation) of the batch under a
    setting; returns its rate, score and value evaluations per placement. Leaves the full beam set."""
    ev.set_beam_widths(setting["widths"])
    if setting["k"]:
        ev.set_pruner(pruner, act_range, k=setting["k"], ply_mask=setting["mask"])
    else:
        ev.set_pruner(None, 0.0)
    ev.stage(w, num_models, seeds, max_moves)
    ev.launch()
    ev.sync()
    t0 = time.perf_counter()
    for _ in range(max(1, steps)):
        ev.launch()
        ev.sync()
    elapsed = (time.perf_counter() - t0) / max(1, steps)
    fit, pl = ev.fetch()
    cnt = ev.fetch_counters()
    ev.set_pruner(None, 0.0)
    ev.set_beam_widths(FULL_WIDTHS)
    placements = int(pl.sum())
    return {"setting": setting["name"], "widths": list(setting["widths"]), "k": setting["k"],
            "ply_mask": setting["mask"], "ms_per_step": elapsed * 1e3, "placements": placements,
            "placements_per_s": placements / elapsed, "mean_score": float(fit.mean()),
            "value_evals_per_placement": float(cnt[:, 1].sum()) / max(placements, 1), "fitness": fit}


def compare_settings(ev: ModelEvaluator, w, num_models: int, seeds, max_moves: int, specs, pruner=None,
                     act_range: float = 0.0, steps: int = 3) -> list:
    """The full beam and every setting of `specs` on the same batch: speed-up over the full beam's placements/s,
    mean-score loss relative to the full beam's |mean| (positive = worse), share of games with identical fitness."""
    full = run_setting(ev, w, num_models, seeds, max_moves, parse_setting("full"), steps=steps)
    out = []
    for spec in specs:
        r = run_setting(ev, w, num_models, seeds, max_moves, parse_setting(spec), pruner, act_range, steps)
        fm = full["mean_score"]
        r["speedup_vs_full"] = r["placements_per_s"] / full["placements_per_s"]
        r["mean_score_loss_pct"] = 100.0 * (fm - r["mean_score"]) / abs(fm) if fm != 0 else 0.0
        r["meets_1pct"] = bool(r["mean_score_loss_pct"] <= 1.0)
        r["same_games"] = float((r.pop("fitness") == full["fitness"]).mean())
        out.append(r)
    full.pop("fitness")
    return [full] + out


def fit_pruner(ev: ModelEvaluator, w, num_models: int, seeds, max_moves: int, every: int, cap_records: int,
               records: int, epochs: int, seed: int = 1):
    """Traces of the full beam on (w, seeds), a sample of `records` of them, training on the GPU and the
    activation-range calibration: (pruner weights, act_range, info)."""
    t0 = time.perf_counter()
    recs, produced, _ = collect_traces(ev, w, num_models, seeds, max_moves, every=every, cap=cap_records)
    rng = np.random.default_rng(seed)
    train = recs if recs.shape[0] <= records else recs[rng.choice(recs.shape[0], records, replace=False)]
    pw, losses, secs = train_pruner(ev, train, epochs=epochs, batch=512, lr=1e-3, seed=seed)
    act = calibrate_act_range(ev, pw, train)
    return pw, act, {"trace_records": int(produced), "kept": int(recs.shape[0]), "trained_on": int(train.shape[0]),
                     "every": every, "epochs": epochs, "train_s": secs, "setup_s": time.perf_counter() - t0,
                     "training_loss": [float(x) for x in losses]}


def _main(argv=None) -> int:
    """Pruner study on a trained model (cfg3 shape): train on the model's own full-beam traces from one seed set,
    then judge every setting against the full beam on `--games` held-out games. Prints one JSON line per
    setting."""
    import argparse
    import json

    from . import MODEL_F16, MODEL_F32W, MODEL_V11, champ_read, expand_seeds

    kinds = {"v11": MODEL_V11, "f16": MODEL_F16, "f32w": MODEL_F32W}
    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--champ", required=True, help=".champ weights of the model (or models: count = n x params)")
    ap.add_argument("--kind", choices=sorted(kinds), default="f16")
    ap.add_argument("--games", type=int, default=1000)
    ap.add_argument("--max-moves", type=int, default=2000)
    ap.add_argument("--trace-games", type=int, default=200)
    ap.add_argument("--every", type=int, default=64)
    ap.add_argument("--records", type=int, default=4_000_000)
    ap.add_argument("--trace-cap", type=int, default=24_000_000,
                    help="trace buffer records; when the games emit more, which ones are kept depends on the order "
                         "of work (200 games at cap 2,000, 1 move in 64, emit ~22.5 M)")
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--garbage", action="store_true")
    ap.add_argument("--settings", default="p8:8,p8:12,p4:12,n8-16-16,n4-8-8,n8-16-16+p8:8",
                    help="comma list: p<k>:<ply mask> pruned, n<k0>-<k1>-<k2> narrower beam, n...+p... both")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0x5EED)
    ap.add_argument("--save", default="", help="write the trained pruner (npy: weights then act_range)")
    ap.add_argument("--load", default="", help="use a pruner written by --save instead of training one")
    args = ap.parse_args(argv)

    kind = kinds[args.kind]
    w = champ_read(args.champ).astype(np.float32)
    ev = ModelEvaluator(model=kind)
    if args.garbage:
        ev.set_garbage(True)
    nm = w.size // ev.num_params
    if args.load:
        saved = np.load(args.load)
        pw, act = saved[:-1].astype(np.float32), float(saved[-1])
        print(json.dumps({"loaded": args.load, "act_range": act}), flush=True)
    else:
        pw, act, info = fit_pruner(ev, w, nm, expand_seeds(args.seed, args.trace_games), args.max_moves,
                                   args.every, args.trace_cap, args.records, args.epochs)
        print(json.dumps(dict(info, act_range=act)), flush=True)
    if args.save:
        np.save(args.save, np.r_[pw, np.float32(act)])
    rows = compare_settings(ev, w, nm, expand_seeds(args.seed + 1, args.games), args.max_moves,
                            args.settings.split(","), pw, act, args.steps)
    for r in rows:
        print(json.dumps(r), flush=True)
    ev.close()
    return 0


if __name__ == "__main__":
    import sys as _sys
    _sys.exit(_main())
