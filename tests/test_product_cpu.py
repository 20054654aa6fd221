"""CPU-side checks of the product library: it loads, exports every entry point include/ta3h.h declares,
its compile-time tables equal the oracle's, and argument errors are reported without touching a GPU."""
import os
import re

import numpy as np
import pytest

import oracle as O
import tetris_ai3_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^\s*(?:const\s+)?\w+[\s\*]+(ta3h_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_entry_point():
    names = declared_symbols()
    assert len(names) >= 15
    L = T.lib()
    for n in names:
        assert hasattr(L, n), n


def test_tables_match_oracle():
    lut, trees = T.tables()
    for p in range(7):
        assert lut[p] == O.lut(p)
    for he in (0, 1):
        assert trees[he] == O.tree(4, bool(he))


def test_seed_expansion_matches_oracle():
    for seed in (0, 1, 0x3E28DF7B1811145B, 2**64 - 1):
        assert list(T.expand_seeds(seed, 7)) == list(O.expand_seeds(seed, 7))


def test_num_params():
    assert T.num_params(T.MODEL_V11) == 97 and T.num_params(T.MODEL_F16) == 1633
    assert T.lib().ta3h_num_params(99) == -1


def test_argument_errors_without_gpu():
    ev = T.ModelEvaluator()
    with pytest.raises(T.HipError, match="NUM_PARAMS"):
        ev.eval(np.zeros(96, dtype=np.float32), 1, [1], 10)
    # zero (model, game) items: an empty result, no launch (model_eval.cu:104-107)
    assert ev.eval(np.zeros(97, dtype=np.float32), 1, np.zeros(0, dtype=np.uint64), 10).size == 0
    with pytest.raises(T.HipError, match="waves_per_game"):
        ev.set_launch(3)
    with pytest.raises(T.HipError, match="unknown model"):
        T.ModelEvaluator(model=7)
    # pruner arguments are checked before any device work
    pw = T.pruner_init(0)
    with pytest.raises(T.HipError, match="k must be"):
        ev.set_pruner(pw, 2.0, k=0)
    with pytest.raises(T.HipError, match="ply_mask"):
        ev.set_pruner(pw, 2.0, k=8, ply_mask=0b10001)
    with pytest.raises(T.HipError, match="act_range"):
        ev.set_pruner(pw, 0.0, k=8)
    with pytest.raises(T.HipError, match="21249"):
        ev.set_pruner(pw[:-1], 2.0)
    ev.set_pruner(None, 0.0)  # back to the full beam: no device work
    for bad in ((0, 32, 32), (17, 32, 32), (16, 33, 32), (16, 32, 0)):
        with pytest.raises(T.HipError, match="set_beam_widths"):
            ev.set_beam_widths(bad)
    ev.set_beam_widths((4, 8, 8))


def test_noalloc_refuses_batches_beyond_its_scratch_without_gpu():
    """ta3h_eval_noalloc checks sizes before any HIP call: a batch whose scratch exceeds the static array is
    refused with TA3H_E_SCRATCH, bad arguments with TA3H_E_ARG, an empty batch succeeds"""
    fake = 1 << 20  # never dereferenced: the checks come first
    assert T.eval_noalloc(T.MODEL_F16, fake, 1000, fake, 1000, 100, fake) == T.E_SCRATCH
    assert T.eval_noalloc(T.MODEL_F32W, fake, 8192, fake, 1, 100, fake) == T.E_SCRATCH  # 348.7 MB of weights
    assert T.eval_noalloc(2, fake, 4, fake, 4, 100, fake) == 1
    assert T.eval_noalloc(T.MODEL_V11, 0, 4, fake, 4, 100, fake) == 1
    assert T.eval_noalloc(T.MODEL_V11, fake, 0, fake, 4, 100, fake) == 0


def test_pruner_quantization_matches_oracle():
    """the library's grid exponents for the integer inference form equal the oracle's over weights and
    activation ranges spanning many binades"""
    import oracle as O
    rng = np.random.default_rng(0)
    for _ in range(100):
        w = rng.normal(0, 10 ** rng.uniform(-3, 2), T.pruner_num_params()).astype(np.float32)
        act = float(10 ** rng.uniform(-3, 3))
        assert T.pruner_quant_info(w, act) == O.pruner_quant_info(w, act)


def test_graph_run_without_prepare_is_a_state_error():
    ev = T.ModelEvaluator()
    ev._graph_models, ev._graph_items = 1, 1
    with pytest.raises(T.HipError, match="no generation graph"):
        ev.graph_run(np.zeros(97, dtype=np.float32))


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 (or a multi-rank launch) exits with a message, before touching
    a device"""
    import subprocess
    import sys
    for args, env in ((["--gpus", "2"], {}), ([], {"WORLD_SIZE": "4"})):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                           timeout=120, env={**os.environ, **env})
        assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_trace_records_canonical_order():
    """training follows the record order, so records are put in an order that depends on their content alone:
    any permutation of a record set gives the same array"""
    from tetris_ai3_amd import learn
    rng = np.random.default_rng(5)
    recs = rng.integers(0, 2**32, size=(5000, 11), dtype=np.uint64).astype(np.uint32)
    recs[100:110] = recs[7]  # duplicates
    a = learn.canonical_order(recs)
    b = learn.canonical_order(recs[rng.permutation(recs.shape[0])])
    assert a.shape == recs.shape and (a == b).all()
    assert sorted(map(bytes, a)) == sorted(map(bytes, recs))
