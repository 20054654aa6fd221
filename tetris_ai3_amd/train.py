"""Training run: CMA-ES islands around the GPU evaluator, one graph replay per generation.

    python -m tetris_ai3_amd.train --kind v11 --iterations 20 --generations 10 --out run/v11

Defaults are the reference trainer's setup (4 islands x 50 population x 50 games, 10 generations per
iteration, README.md:9) with sigma 0.3 and island means ~ N(0, 0.5^2). Prints one JSON line per iteration
(best / champion fitness, generation latency, placements per second) and writes the champion as
<out>_<iteration>.champ (u64 count + doubles, the reference's .champ format) and <out>.champ, and the whole
trainer state as <out>.ckpt; `--resume <out>.ckpt` continues a run exactly where it stopped (the checkpoint's
configuration replaces the command line's).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import MODEL_F16, MODEL_F32W, MODEL_V11, ModelEvaluator, Trainer, autotune, champ_write, expand_seeds

KINDS = {"v11": MODEL_V11, "f16": MODEL_F16, "f32w": MODEL_F32W}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--kind", choices=sorted(KINDS), default="v11")
    ap.add_argument("--islands", type=int, default=4)
    ap.add_argument("--population", type=int, default=50)
    ap.add_argument("--games", type=int, default=50)
    ap.add_argument("--max-moves", type=int, default=2000)
    ap.add_argument("--generations", type=int, default=10)
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--sigma0", type=float, default=0.3)
    ap.add_argument("--init-std", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=0x7E7215)
    ap.add_argument("--garbage", action="store_true", help="start games with 0-4 garbage rows")
    ap.add_argument("--out", default="", help="path prefix of the .champ files and the .ckpt checkpoint; empty = none")
    ap.add_argument("--resume", default="", help="continue from a .ckpt checkpoint")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the network's default launch shape instead of tuning it on a first-generation batch")
    args = ap.parse_args(argv)

    if args.resume:
        trainer = Trainer.load(args.resume)
        kind = trainer.kind
        args.islands, args.population = trainer.cfg_islands, trainer.cfg_population
        args.games, args.generations = trainer.cfg_games, trainer.cfg_generations
    else:
        kind = KINDS[args.kind]
        trainer = Trainer(kind=kind, islands=args.islands, population=args.population, games=args.games,
                          max_moves=args.max_moves, generations=args.generations, sigma0=args.sigma0,
                          init_std=args.init_std, seed=args.seed)
    ev = ModelEvaluator(model=kind)
    if args.garbage:
        ev.set_garbage(True)
    if args.out:
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    if not args.no_autotune:
        # launch shape and first-phase budget tuned on a batch shaped like a generation: island means ~ N(0, init_std^2)
        # plus sigma0 noise, the seeds of a generation (results never depend on the shape)
        M = args.islands * args.population
        rng = np.random.default_rng(args.seed)
        w0 = (rng.normal(0.0, args.init_std, (args.islands, 1, trainer.num_params))
              + args.sigma0 * rng.normal(0.0, 1.0, (args.islands, args.population, trainer.num_params)))
        tuned = autotune(ev, w0.astype(np.float32).reshape(-1), M, expand_seeds(args.seed, args.games),
                         trainer.cfg_max_moves)
        print(json.dumps({"autotune": tuned}), flush=True)
    ev._items = args.islands * args.population * args.games  # the trainer's generation batch
    for it in range(args.iterations):
        t0 = time.perf_counter()
        rep = trainer.run_iteration(evaluator=ev)
        wall = time.perf_counter() - t0
        _, placements = ev.fetch()  # the last generation's per-item placement counts
        w, fit = trainer.champion()
        line = {"iteration": rep["iteration"], "best_fitness": rep["best_fitness"],
                "champion_fitness": fit, "last_mean_fitness": rep["last_mean_fitness"],
                "generation_ms": rep["device_ms"] / args.generations, "wall_s": wall,
                "last_generation_placements": int(placements.sum()),
                "last_generation_game_p50": float(np.percentile(placements, 50)),
                "last_generation_game_p99": float(np.percentile(placements, 99))}
        print(json.dumps(line), flush=True)
        if args.out:
            champ_write(f"{args.out}_{rep['iteration']:03d}.champ", w)
            champ_write(f"{args.out}.champ", w)
            trainer.save(f"{args.out}.ckpt")
    ev.close()
    trainer.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
