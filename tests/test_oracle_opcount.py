"""The roofline's op model (bench.py C_FIT, C_PLACE_CLEAR, C_FEAT, C_MLP) is frozen from the instrumented scalar
restatement of one candidate (oracle/opcount.cpp): every scalar operation of fit, place + clear, the features and
the value network counted over the legal candidates of real full-beam games. The restatement computes the same
drops, boards, features and values as the oracle, and the frozen constants are its rounded averages on this
fixed sample."""
import numpy as np
import pytest

import bench
import oracle as O


def sample(kind):
    w, M, seeds = bench.workload(kind)
    P = O.num_params(kind)
    recs = np.concatenate([O.trace_records(kind, w[m * P:(m + 1) * P], int(s), 150, 7)
                           for m in range(2) for s in seeds[:2]])
    return w[:P], recs


def _hist(cleared):
    return 0 if cleared == 0 else 1 << (8 * (cleared - 1))


@pytest.mark.parametrize("kind", [O.MODEL_V11, O.MODEL_F16])
def test_op_counts_restate_the_oracle_and_freeze_the_op_model(kind):
    w, recs = sample(kind)
    n = recs.shape[0]
    assert n > 20000
    ops, drop, boards, feats, vals = O.candidate_ops(kind, w, recs)
    rng = np.random.default_rng(5)
    nf = 10 if kind == O.MODEL_V11 else 16
    for r in rng.choice(n, 400, replace=False):  # the restatement computes what the oracle computes
        meta = int(recs[r, 10])
        piece, o, x = meta & 7, (meta >> 3) & 3, ((meta >> 5) & 15) - 2
        b = O.Board(recs[r, :10])
        legal, d = b.fit(piece, o, x)
        assert legal and drop[r] == d
        b.place(piece, o, x, d)
        cleared = b.clear_lines()
        assert np.array_equal(boards[r], b.cols)
        want = O.features(kind, _hist(cleared), b.cols, False)
        assert feats[r, :nf].tobytes() == want.tobytes()
        assert np.float32(vals[r]) == np.float32(O.model_value(kind, w, _hist(cleared), b.cols, False))
    per = ops.astype(np.float64) / n
    assert round(per[0]) == bench.C_FIT
    assert round(per[1]) == bench.C_PLACE_CLEAR[kind]
    assert per[2] == bench.C_FEAT[kind]  # data-independent
    assert per[3] == bench.C_MLP[kind]
