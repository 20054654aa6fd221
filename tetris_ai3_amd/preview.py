"""Preview: watch models play, the counterpart of the reference's AiPreview (ta3_ai/src/preview.cpp:40-94) for a
machine without a display.

    python -m tetris_ai3_amd.preview --champ model.champ [--kind v11] [--games 4] [--moves 300]

The games run on the GPU with the full beam (one game per model of the .champ file, or --games games of its one
model on seeds drawn from --seed), with the move log on; the boards they pass through are rebuilt by the host
replay (ta3h_replay) and drawn side by side as text, 20 visible rows each (the reference's renderer hides the three
spawn rows, ta3_sim/incl/ta3/sim/ui/ai_renderer2.hpp:28-29). Frames are animated in the terminal every --delay
seconds, or written to --out as plain text, one frame every --every moves.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import MODEL_F16, MODEL_F32W, MODEL_V11, ModelEvaluator, champ_read, expand_seeds, replay

KINDS = {"v11": MODEL_V11, "f16": MODEL_F16, "f32w": MODEL_F32W}
PIECES = "IOTSZJL"
VISIBLE = 20  # rows 3..22 of the 23


def draw_board(cols) -> list[str]:
    rows = []
    for y in range(23 - VISIBLE, 23):
        rows.append("|" + "".join("#" if (int(cols[x]) >> y) & 1 else " " for x in range(10)) + "|")
    rows.append("+" + "-" * 10 + "+")
    return rows


def frame(games, i: int) -> str:
    """the boards of every game after move i (a game that ended shows its last board)"""
    blocks, heads = [], []
    for g in games:
        n = len(g["boards"])
        j = min(i, n - 1)
        cols = g["boards"][j] if n else np.zeros(10, dtype=np.uint32)
        state = "top-out" if g["ended"] and i >= n - 1 else f"{PIECES[g['pieces'][j]]} placed" if n else ""
        heads.append(f"game {g['index']:<3} move {min(i, n - 1) + 1:>4}".ljust(12))
        heads.append(f"lines {int(g['lines'][: j + 1].sum()):>4} {state}".ljust(12))
        blocks.append(draw_board(cols))
    lines = ["   ".join(h.ljust(12) for h in heads[0::2]), "   ".join(h.ljust(12) for h in heads[1::2])]
    for r in range(len(blocks[0])):
        lines.append("   ".join(b[r].ljust(12) for b in blocks))
    return "\n".join(lines)


def play(ev: ModelEvaluator, w, num_models: int, seeds, max_moves: int, garbage: bool):
    """the games on the GPU (move log on), then their boards by host replay"""
    ev.set_move_log(True)
    if garbage:
        ev.set_garbage(True)
    ev.stage(w, num_models, seeds, max_moves)
    ev.launch()
    fit, pl = ev.fetch()
    log = ev.fetch_moves()
    games = []
    for item in range(num_models * len(seeds)):
        seed = int(seeds[item % len(seeds)])
        boards, pieces, cleared = replay(seed, log[item, :pl[item]], garbage)
        games.append({"index": item, "boards": boards, "pieces": pieces, "lines": cleared.astype(np.int64),
                      "fitness": float(fit[item]), "placements": int(pl[item]),
                      "ended": int(pl[item]) < max_moves})
    return games


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--champ", required=True, help=".champ weights (one model, or several back to back)")
    ap.add_argument("--kind", choices=sorted(KINDS), default="v11")
    ap.add_argument("--games", type=int, default=4, help="games of a one-model file")
    ap.add_argument("--moves", type=int, default=300)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--garbage", action="store_true")
    ap.add_argument("--delay", type=float, default=0.05, help="seconds between frames (the reference's 50 ms)")
    ap.add_argument("--every", type=int, default=1, help="draw every n-th move")
    ap.add_argument("--out", default="", help="write the frames to this text file instead of animating")
    args = ap.parse_args(argv)

    kind = KINDS[args.kind]
    ev = ModelEvaluator(model=kind)
    w = champ_read(args.champ).astype(np.float32)
    nm = w.size // ev.num_params
    if nm * ev.num_params != w.size or nm == 0:
        raise SystemExit(f"{args.champ}: {w.size} weights is not a whole number of {ev.num_params}-weight models")
    seeds = expand_seeds(args.seed or int(time.time_ns()), 1 if nm > 1 else args.games)
    games = play(ev, w, nm, seeds, args.moves, args.garbage)
    ev.close()
    longest = max(len(g["boards"]) for g in games)
    out = open(args.out, "w") if args.out else None
    for i in list(range(0, longest, max(1, args.every))) + [longest - 1]:
        text = frame(games, i)
        if out:
            out.write(text + "\n\n")
        else:
            sys.stdout.write("\x1b[H\x1b[2J" + text + "\n")
            sys.stdout.flush()
            time.sleep(args.delay)
    summary = "  ".join(f"game {g['index']}: {g['placements']} pieces, fitness {g['fitness']:.1f}" for g in games)
    if out:
        out.write(summary + "\n")
        out.close()
    print(summary)
    return 0


if __name__ == "__main__":
    sys.exit(main())
