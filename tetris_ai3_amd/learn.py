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


def _main(argv=None) -> int:
    """Pruner study on a trained model (cfg3 shape): train on the model's own full-beam traces from one seed
    set, then judge pruned vs full beam on `--games` fresh games. Prints one JSON line per setting."""
    import argparse
    import json

    from . import MODEL_F16, MODEL_F32W, MODEL_V11, champ_read, expand_seeds

    kinds = {"v11": MODEL_V11, "f16": MODEL_F16, "f32w": MODEL_F32W}
    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--champ", required=True, help=".champ weights of the model (or models: count = n x params)")
    ap.add_argument("--kind", choices=sorted(kinds), default="f16")
    ap.add_argument("--games", type=int, default=10000)
    ap.add_argument("--max-moves", type=int, default=2000)
    ap.add_argument("--trace-games", type=int, default=200)
    ap.add_argument("--every", type=int, default=8)
    ap.add_argument("--records", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--settings", default="8:12,8:14,4:12,4:14",
                    help="comma list of k:ply_mask (mask bits 1, 2 = middle levels, 3 = leaves)")
    ap.add_argument("--seed", type=int, default=0x5EED)
    ap.add_argument("--save", default="", help="write the trained pruner (npy: weights then act_range)")
    args = ap.parse_args(argv)

    kind = kinds[args.kind]
    w = champ_read(args.champ).astype(np.float32)
    ev = ModelEvaluator(model=kind)
    nm = w.size // ev.num_params
    trace_seeds = expand_seeds(args.seed, args.trace_games)
    recs, produced, _ = collect_traces(ev, w, nm, trace_seeds, args.max_moves, args.every, args.records * 2)
    rng = np.random.default_rng(args.seed)
    if recs.shape[0] > args.records:
        recs = recs[rng.choice(recs.shape[0], args.records, replace=False)]
    pw, losses, secs = train_pruner(ev, recs, epochs=args.epochs, seed=1)
    act = calibrate_act_range(ev, pw, recs)
    print(json.dumps({"traces": int(produced), "trained_on": int(recs.shape[0]), "train_s": secs,
                      "epoch_loss": [float(x) for x in losses], "act_range": act,
                      "label_stats": label_stats(recs)}), flush=True)
    if args.save:
        np.save(args.save, np.r_[pw, np.float32(act)])
    seeds = expand_seeds(args.seed + 1, args.games)

    def run():
        ev.stage(w, nm, seeds, args.max_moves)
        ev.launch()
        ev.sync()
        ms = ev.last_kernel_ms()
        fit, pl = ev.fetch()
        cnt = ev.fetch_counters()
        return fit, pl, cnt, ms

    ev.set_pruner(None, 0.0)
    full, full_pl, full_cnt, full_ms = run()
    base = {"beam": "full", "mean_score": float(full.mean()), "placements": int(full_pl.sum()), "ms": full_ms,
            "placements_per_s": float(full_pl.sum()) / full_ms * 1e3,
            "value_evals_per_placement": float(full_cnt[:, 1].sum()) / float(full_pl.sum())}
    print(json.dumps(base), flush=True)
    for item in args.settings.split(","):
        k, mask = (int(v) for v in item.split(":"))
        ev.set_pruner(pw, act, k=k, ply_mask=mask)
        fit, pl, cnt, ms = run()
        rate = float(pl.sum()) / ms * 1e3
        print(json.dumps({"beam": "pruned", "k": k, "ply_mask": mask, "mean_score": float(fit.mean()),
                          "mean_score_loss_pct": 100.0 * (1.0 - float(fit.mean()) / float(full.mean())),
                          "same_games": float((fit == full).mean()), "placements": int(pl.sum()), "ms": ms,
                          "placements_per_s": rate, "speedup": rate / base["placements_per_s"],
                          "value_evals_per_placement": float(cnt[:, 1].sum()) / float(pl.sum())}), flush=True)
    ev.close()
    return 0


if __name__ == "__main__":
    import sys as _sys
    _sys.exit(_main())
