"""GPU: the CMA-ES island loop evaluated on the device equals the identical loop evaluated by the CPU oracle
(same seeds, same host optimiser): identical reports, champion and island states, because every game is
bit-exact. Each generation runs as one graph replay."""
import sys

import numpy as np
import pytest

import oracle as O
import tetris_ai3_amd as T

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", [O.MODEL_V11, O.MODEL_F16])
def test_device_training_equals_cpu_training(kind):
    say(f"case: trainer kind={kind} 2 islands x 8 pop x 4 games, 2 iterations x 3 generations")
    cfg = dict(kind=kind, islands=2, population=8, games=4, max_moves=60, generations=3, seed=21)
    gpu, cpu = T.Trainer(**cfg), T.Trainer(**cfg)
    ev = T.ModelEvaluator(model=kind)
    host = lambda w, m, seeds, max_moves: O.eval_host_mt(kind, w, m, seeds, max_moves, threads=8)[0]
    for _ in range(2):
        rg = gpu.run_iteration(evaluator=ev)
        rc = cpu.run_iteration(evaluate=host)
        assert rg["device_ms"] > 0
        rg.pop("device_ms"), rc.pop("device_ms")
        assert rg == rc
    wg, fg = gpu.champion()
    wc, fc = cpu.champion()
    assert fg == fc and (wg == wc).all()
    for i in range(2):
        mg, sg = gpu.island_state(i)
        mc, sc = cpu.island_state(i)
        assert sg == sc and (mg == mc).all()
    ev.close()


def test_training_command_resumes_from_its_checkpoint(tmp_path, capsys):
    """python -m tetris_ai3_amd.train: one iteration, then --resume from the .ckpt it wrote for one more, gives
    the champion of an uninterrupted two-iteration run"""
    from tetris_ai3_amd import train
    say("case: training command, checkpoint and resume")
    common = ["--kind", "v11", "--islands", "2", "--population", "6", "--games", "3", "--max-moves", "50",
              "--generations", "2", "--seed", "33"]
    whole, part = str(tmp_path / "whole"), str(tmp_path / "part")
    assert train.main(common + ["--iterations", "2", "--out", whole]) == 0
    assert train.main(common + ["--iterations", "1", "--out", part]) == 0
    assert train.main(["--resume", part + ".ckpt", "--iterations", "1", "--out", part]) == 0
    capsys.readouterr()
    assert open(whole + ".champ", "rb").read() == open(part + ".champ", "rb").read()
    assert open(whole + ".ckpt", "rb").read() == open(part + ".ckpt", "rb").read()
