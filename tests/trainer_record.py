"""Training-loop record (configs[2] shape): the CMA-ES island loop on the GPU against the identically seeded loop on
the CPU oracle.

    python tests/trainer_record.py compare --out <json>   # GPU vs CPU, same seeds: one iteration of 10 generations

The full-size configuration (4 islands x 50 population x 50 games, cap 2,000, 10 generations per iteration) runs
through the training command, which prints per-iteration generation latency and placements:
    python -m tetris_ai3_amd.train --kind v11 --iterations 20
A CPU generation at that size is ~10^4 core-seconds, so the CPU comparison runs the same loop at 4 islands x 8 x 8
games, cap 300 (the reference runner's cap, ta3_trainer/runner/src/main.cpp:29-30): the loop is identical, only
the batch is smaller. Test infrastructure: it loads the oracle as the second evaluator.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
import tetris_ai3_amd as T  # noqa: E402


def compare(out: str, iterations: int, threads: int) -> int:
    cfg = dict(kind=T.MODEL_V11, islands=4, population=8, games=8, max_moves=300, generations=10, seed=0x7E7215)
    gpu, cpu = T.Trainer(**cfg), T.Trainer(**cfg)
    ev = T.ModelEvaluator(model=T.MODEL_V11)

    def host(w, m, seeds, max_moves):
        return O.eval_host_mt(O.MODEL_V11, w, m, seeds, max_moves, threads=threads)[0]

    rec = {"config": cfg, "cpu_threads": threads, "iterations": []}
    for it in range(iterations):
        t0 = time.perf_counter()
        rg = gpu.run_iteration(evaluator=ev)
        tg = time.perf_counter() - t0
        t0 = time.perf_counter()
        rc = cpu.run_iteration(evaluate=host)
        tc = time.perf_counter() - t0
        wg, fg = gpu.champion()
        wc, fc = cpu.champion()
        same = {k: rg[k] == rc[k] for k in rg if k != "device_ms"}
        line = {"iteration": it + 1, "gpu": rg, "cpu": rc, "gpu_wall_s": tg, "cpu_wall_s": tc,
                "gpu_champion_fitness": fg, "cpu_champion_fitness": fc,
                "champions_identical": bool(fg == fc and np.array_equal(wg, wc)),
                "reports_identical": all(same.values())}
        rec["iterations"].append(line)
        print(json.dumps(line), flush=True)
    ev.close()
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    return 0 if all(x["champions_identical"] and x["reports_identical"] for x in rec["iterations"]) else 1


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("mode", choices=["compare"])
    ap.add_argument("--out", required=True)
    ap.add_argument("--iterations", type=int, default=1)
    ap.add_argument("--threads", type=int, default=min(os.cpu_count() or 1, 16))
    args = ap.parse_args()
    return compare(args.out, args.iterations, args.threads)


if __name__ == "__main__":
    sys.exit(main())
