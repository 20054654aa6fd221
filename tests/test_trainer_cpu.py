"""The CMA-ES island loop (host code of the library) and .champ files, on the CPU.

- sep-CMA-ES converges linearly on the sphere and on a 10^6-conditioned separable ellipsoid (the reference's
  own optimiser tests check convergence on sphere / ellipsoids, dd_cmaes_test.cpp:39-52).
- Runs are deterministic per seed (dd_cmaes_test.cpp:58-64) and differ across seeds.
- The training loop on the CPU oracle: model fitness is the mean of its games, the champion is the best model.
- .champ: u64 count + doubles (trainer2.cpp:251-265); round trip, and the reference's shipped v11 champion
  reads back as the golden weights when the reference checkout is present.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import tetris_ai3_amd as T

HERE = os.path.dirname(os.path.abspath(__file__))
REF_CHAMP = "/root/reference/models/v11_16_32_32_score_296.champ"


def quadratic(target, scale):
    def f(w, m, seeds, max_moves):
        W = w.reshape(m, -1).astype(np.float64)
        return np.repeat(-(scale * (W - target) ** 2).sum(1), len(seeds)).astype(np.float32)
    return f


def test_sep_cmaes_converges_on_sphere():
    target = np.random.default_rng(0).normal(0, 1, 97)
    t = T.Trainer(kind=T.MODEL_V11, islands=1, population=20, games=1, max_moves=1, generations=10, seed=1)
    first = t.run_iteration(evaluate=quadratic(target, 1.0))["best_fitness"]
    for _ in range(49):
        last = t.run_iteration(evaluate=quadratic(target, 1.0))["best_fitness"]
    assert first < -1.0 and last > -1e-6
    mean, sigma = t.island_state(0)
    assert np.abs(mean - target).max() < 1e-2 and sigma < 1e-3


def test_sep_cmaes_converges_on_separable_ellipsoid():
    scale = 10.0 ** (6 * np.arange(97) / 96)
    t = T.Trainer(kind=T.MODEL_V11, islands=1, population=20, games=1, max_moves=1, generations=10, seed=2)
    first = t.run_iteration(evaluate=quadratic(0.0, scale))["best_fitness"]
    for _ in range(99):
        last = t.run_iteration(evaluate=quadratic(0.0, scale))["best_fitness"]
    assert last > first * 1e-5


def oracle_eval(kind):
    return lambda w, m, seeds, max_moves: O.eval_host_mt(kind, w, m, seeds, max_moves, threads=8)[0]


def small_trainer(seed):
    return T.Trainer(kind=T.MODEL_V11, islands=2, population=6, games=3, max_moves=40, generations=2, seed=seed)


def test_training_loop_is_deterministic_per_seed():
    runs = []
    for seed in (7, 7, 8):
        t = small_trainer(seed)
        reps = [t.run_iteration(evaluate=oracle_eval(O.MODEL_V11)) for _ in range(2)]
        runs.append((reps, t.champion()))
    assert runs[0][0] == runs[1][0] and (runs[0][1][0] == runs[1][1][0]).all()
    assert runs[0][0] != runs[2][0]


def test_champion_fitness_is_its_mean_game_fitness():
    t = small_trainer(9)
    rep = t.run_iteration(evaluate=oracle_eval(O.MODEL_V11))
    w, fit = t.champion()
    assert fit == rep["champion_fitness"] == rep["best_fitness"]
    # replay the champion on the iteration's seeds (simulate()'s expansion of the iteration seed)
    seeds = O.expand_seeds(rep["iteration_seed"], 3)
    games, _ = O.eval_host_mt(O.MODEL_V11, w.astype(np.float32), 1, seeds, 40)
    assert fit == sum(float(g) for g in games) / 3


def test_champ_round_trip(tmp_path):
    w = np.random.default_rng(3).normal(0, 100, 97)
    p = str(tmp_path / "m.champ")
    T.champ_write(p, w)
    raw = open(p, "rb").read()
    assert len(raw) == 8 + 97 * 8 and int.from_bytes(raw[:8], "little") == 97
    assert (T.champ_read(p) == w).all()
    with open(p, "ab") as f:
        f.write(b"x")
    with pytest.raises(T.HipError, match="malformed"):
        T.champ_read(p)


@pytest.mark.skipif(not os.path.exists(REF_CHAMP), reason="reference checkout not present")
def test_reference_champion_file_reads_as_the_golden_weights(tmp_path):
    w = T.champ_read(REF_CHAMP)
    with open(os.path.join(HERE, "golden", "v11_champion.json")) as f:
        golden = np.array(json.load(f)["weights_f64"], dtype=np.float64)
    assert w.size == 97 and (w == golden).all()
    out = str(tmp_path / "copy.champ")
    T.champ_write(out, w)
    assert open(out, "rb").read() == open(REF_CHAMP, "rb").read()


def test_checkpoint_resumes_identically(tmp_path):
    """a trainer saved between iterations and loaded again continues exactly as the uninterrupted one (the
    reference checks that a serialized optimiser resumes identically, dd_cmaes_test.cpp:82-104)"""
    ev = oracle_eval(O.MODEL_V11)
    t = small_trainer(11)
    for _ in range(2):
        t.run_iteration(evaluate=ev)
    path = str(tmp_path / "trainer.ckpt")
    t.save(path)
    cont = [t.run_iteration(evaluate=ev) for _ in range(2)]
    w_cont, f_cont = t.champion()
    u = T.Trainer.load(path)
    assert u.num_params == 97
    res = [u.run_iteration(evaluate=ev) for _ in range(2)]
    w_res, f_res = u.champion()
    assert res == cont and f_res == f_cont and (w_res == w_cont).all()
    for i in range(2):
        (ma, sa), (mb, sb) = t.island_state(i), u.island_state(i)
        assert sa == sb and (ma == mb).all()
    t.close()
    u.close()


def test_checkpoint_rejects_damaged_files(tmp_path):
    t = small_trainer(12)
    t.run_iteration(evaluate=oracle_eval(O.MODEL_V11))
    path = str(tmp_path / "trainer.ckpt")
    t.save(path)
    raw = open(path, "rb").read()
    for name, data in (("truncated", raw[:-5]), ("trailing", raw + b"\0"), ("magic", b"X" + raw[1:]),
                       ("empty", b"")):
        bad = str(tmp_path / f"{name}.ckpt")
        with open(bad, "wb") as f:
            f.write(data)
        with pytest.raises(T.HipError, match="cannot load"):
            T.Trainer.load(bad)
    with pytest.raises(T.HipError, match="cannot load"):
        T.Trainer.load(str(tmp_path / "missing.ckpt"))
    assert T.Trainer.load(path).num_params == 97
