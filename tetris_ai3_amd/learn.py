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


def collect_traces(ev: ModelEvaluator, weights, num_models: int, seeds, max_moves: int, every: int, cap: int):
    """Plays the batch with the full beam and returns (records uint32 [n, 11], produced, fitness)."""
    ev.set_trace(cap, every)
    ev.stage(weights, num_models, seeds, max_moves)
    ev.launch()
    fit, _ = ev.fetch()
    recs, produced = ev.fetch_trace()
    ev.set_trace(0)
    return recs, produced, fit


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


def list_recall(ev: ModelEvaluator, w, act_range: float, recs, k: int) -> dict:
    """Per ply: the fraction of positive records that rank in the top k of their list by pruner logit
    (a list = the records of one parent board and piece at one ply, in record order)."""
    logit = ev.pruner_logits(w, act_range, recs)
    meta = recs[:, 10]
    ply = (meta >> PLY_SHIFT) & 3
    lab = (meta >> LABEL_SHIFT) & 1
    key = np.concatenate([recs[:, :10], (meta & 7)[:, None], ply[:, None]], axis=1)
    _, grp = np.unique(key, axis=0, return_inverse=True)
    grp = grp.reshape(-1)
    # rank within the group by logit descending, record order ascending
    order = np.lexsort((np.arange(len(logit)), -logit, grp))
    rank = np.empty(len(order), dtype=np.int64)
    starts = np.r_[0, np.flatnonzero(np.diff(grp[order])) + 1]
    sizes = np.diff(np.r_[starts, len(order)])
    rank[order] = np.arange(len(order)) - np.repeat(starts, sizes)
    out = {}
    for p in (1, 2, 3):
        m = (ply == p) & (lab == 1)
        out[p] = float((rank[m] < k).mean()) if m.any() else float("nan")
    return out
