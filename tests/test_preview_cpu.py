"""The preview's host replay (ta3h_replay) rebuilds the boards of a move log exactly as the oracle's games pass
through them (with and without garbage rows), and the text renderer draws them (no GPU)."""
import numpy as np
import pytest

import oracle as O
import tetris_ai3_amd as T
from tetris_ai3_amd import preview


def encode(moves):
    m = moves.astype(np.uint32)
    return ((m[:, 0] << 7) | (m[:, 1] << 5) | m[:, 2]).astype(np.uint8)


@pytest.mark.parametrize("garbage", [False, True])
def test_replay_rebuilds_oracle_boards(garbage):
    rng = np.random.default_rng(31)
    for kind in (O.MODEL_V11, O.MODEL_F16):
        w = rng.normal(0, 0.5, O.num_params(kind)).astype(np.float32)
        for seed in (3, 77, 1234):
            if garbage:
                with O.garbage():
                    moves, boards, _ = O.play_trace(kind, w, seed, 150)
            else:
                moves, boards, _ = O.play_trace(kind, w, seed, 150)
            got, pieces, cleared = T.replay(seed, encode(moves), garbage)
            assert got.shape[0] == moves.shape[0]
            assert np.array_equal(got, boards)
            alive = moves[:, 3] != 255
            assert np.array_equal(cleared[alive], moves[alive, 3])
            assert set(np.unique(pieces)) <= set(range(7))


def test_replay_stops_at_top_out_and_rejects_impossible_moves():
    w = np.random.default_rng(0).uniform(-2, 2, 97).astype(np.float32)  # a random network that tops out early
    moves, boards, _ = O.play_trace(O.MODEL_V11, w, 9, 400)
    assert moves.shape[0] < 400 and moves[-1, 3] == 255
    log = encode(moves)
    got, _, _ = T.replay(9, np.r_[log, log[:5]])  # moves past the top-out are not replayed
    assert got.shape[0] == moves.shape[0] and np.array_equal(got, boards)
    with pytest.raises(T.HipError):
        T.replay(9, np.array([0x80 | (0 << 5) | 30], dtype=np.uint8))  # column 28: off the board


def test_frames_draw_every_board():
    moves, _, _ = O.play_trace(O.MODEL_V11, np.random.default_rng(2).normal(0, .5, 97).astype(np.float32), 5, 40)
    boards, pieces, cleared = T.replay(5, encode(moves))
    g = {"index": 0, "boards": boards, "pieces": pieces, "lines": cleared.astype(np.int64), "ended": False}
    text = preview.frame([g, dict(g, index=1)], 10)
    rows = text.split("\n")
    assert len(rows) == 2 + 20 + 1
    assert rows[2 + 19].count("|") == 4  # two boards side by side
    filled = sum(r.count("#") for r in rows[2:]) // 2
    assert filled == sum(bin(int(c) & ((1 << 23) - 1) & ~0b111).count("1") for c in boards[10])
