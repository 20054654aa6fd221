"""Pins the oracle's board, engine and RNG against the reference's own known answers.

Known answers come from the reference tests:
  ta3_sim/tests/src/board2_test.cpp, ta3_sim/tests/src/tetris_engine_test.cpp,
and from published test vectors of xoshiro256** / splitmix64.
"""
import numpy as np
import pytest

import oracle as O

I, OP, T, S, Z, J, L = range(7)
TOP, RIGHT, BOTTOM, LEFT = range(4)
ROWS = 23


# --- RNG: published vectors ------------------------------------------------------------------------
def test_xoshiro256ss_published_vector():
    # xoshiro256** reference output for the state {1, 2, 3, 4}
    assert O.rng_outputs_from_state([1, 2, 3, 4], 4) == [11520, 0, 1509978240, 1215971899390074240]


def test_splitmix64_seeding_published_vector():
    # seed(s) fills the state with successive splitmix64 outputs (xoshiro256ss.hpp:20-28);
    # splitmix64 from 0 is the published e220a8397b1dcdaf, 6e789e6aa1b965f4, 06c45d188009454f, f88bb8a8724c81ec.
    import ctypes as C
    st = (C.c_uint64 * 4)()
    O.lib().ta3o_rng_seed(0, st)
    assert list(st) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]


def test_simulate_seed_expansion_is_the_rng_stream():
    # gpu::simulate: seeds[g] = xoshiro256ss{seed}() in order (model_eval.hpp:92-95)
    assert list(O.expand_seeds(77, 5)) == O.rng_outputs(77, 5)


# --- board2_test.cpp known answers -------------------------------------------------------------------
def fill_row(b, row):  # board2_test.cpp:12-16
    b.place(I, TOP, 0, row - 1)
    b.place(I, TOP, 4, row - 1)
    b.place(OP, TOP, 7, row - 1)


def fill_rows(b, n):  # board2_test.cpp:19-27
    bottom = ROWS - 1
    for row in range(bottom, bottom - n, -1):
        b.place(I, TOP, 0, row - 1)
        b.place(I, TOP, 4, row - 1)
    for row in range(bottom, bottom - n, -2):
        b.place(OP, TOP, 7, row - 1)


def build_overhang(b):  # board2_test.cpp:31-36
    b.drop_place(OP, TOP, 2, 0)
    b.drop_place(OP, TOP, 2, 0)
    b.drop_place(OP, TOP, 2, 0)
    b.drop_place(I, TOP, 3, 0)


def test_drops_and_clears():  # :38-46
    b = O.Board()
    fill_row(b, 22)
    assert b.full_lines() == 1
    assert b.clear_lines() == 1
    assert b.full_lines() == 0
    assert b.drop_location(OP, TOP, 7, 0) == 20


def test_availability():  # :48-54, :65-95
    b = O.Board()
    b.place(OP, TOP, 3, 21)
    assert not b.available(OP, TOP, 3, 21)
    assert b.available(OP, TOP, 3, 19)
    e = O.Board()
    assert e.available(OP, TOP, 3, 0) and e.available(I, TOP, 0, 0) and e.available(I, RIGHT, 0, 0)
    assert e.available(OP, TOP, -1, 0) and not e.available(OP, TOP, -2, 0)
    assert e.available(OP, TOP, 7, 0) and not e.available(OP, TOP, 8, 0)
    assert e.available(OP, TOP, 3, 21) and not e.available(OP, TOP, 3, 22)


def test_drop_locations():  # :97-121
    e = O.Board()
    assert e.drop_location(OP, TOP, 3, 0) == 21
    assert e.drop_location(I, RIGHT, 0, 0) == 19
    b = O.Board()
    b.place(OP, TOP, 3, 21)
    assert b.drop_location(OP, TOP, 3, 0) == 19


def test_overhang_holes_and_tuck():  # :123-133
    b = O.Board()
    build_overhang(b)
    assert b.holes() == 12
    assert b.available(OP, TOP, 4, 17)
    assert b.drop_location(OP, TOP, 4, 17) == 21


def test_every_piece_lands_and_stacks():  # :135-154
    for t in range(7):
        for o in range(4):
            b = O.Board()
            assert b.available(t, o, 3, 0)
            first = b.drop_location(t, o, 3, 0)
            assert b.available(t, o, 3, first)
            b.place(t, o, 3, first)
            assert not b.available(t, o, 3, first)
            assert b.drop_location(t, o, 3, 0) < first


def test_clears():  # :189-256
    b = O.Board()
    b.place(OP, TOP, 3, 21)
    assert b.full_lines() == 0 and b.clear_lines() == 0
    for n in (2, 4):
        b = O.Board()
        fill_rows(b, n)
        assert b.full_lines() == n
        assert b.clear_lines() == n
        assert b.full_lines() == 0
        assert b.holes() == 0
    b = O.Board()
    fill_rows(b, 2)
    b.drop_place(OP, TOP, 0, 0)
    assert b.clear_lines() == 2 and b.full_lines() == 0 and b.holes() == 0
    b = O.Board()
    fill_row(b, 22)
    assert b.clear_lines() == 1
    assert b.drop_location(OP, TOP, 7, 0) == 20


def test_holes():  # :258-269
    b = O.Board()
    b.place(OP, TOP, 0, 20)
    assert b.holes() == 2
    assert O.Board().holes() == 0


# --- brute-force cell-grid restatement of the board ops on random boards ------------------------------
def grid_of(cols):
    return [[(int(cols[x]) >> y) & 1 for y in range(ROWS)] for x in range(10)]


def cols_of(grid):
    return np.array([sum(grid[x][y] << y for y in range(ROWS)) for x in range(10)], dtype=np.uint32)


def piece_cells(t, o, ox, oy):
    cols, left = O.piece_shape(t, o)
    cells = []
    for c in range(4):
        nib = (cols >> (4 * c)) & 0xF
        for y in range(4):
            if (nib >> y) & 1:
                cells.append((ox + left + c, oy + y))
    return cells


def brute_drop(grid, t, o, x):
    """cell-by-cell hard drop from y=0: legal iff the spawn cells are free (columns in range);
    then fall while every cell below stays free and in the field."""
    cells = piece_cells(t, o, x, 0)
    if any(g[1] >= ROWS or grid[g[0]][g[1]] for g in cells):
        return None
    d = 0
    while all(c[1] + d + 1 < ROWS and not grid[c[0]][c[1] + d + 1] for c in cells):
        d += 1
    return d


def brute_clear(grid):
    full = [y for y in range(ROWS) if all(grid[x][y] for x in range(10))]
    out = [[0] * ROWS for _ in range(10)]
    for x in range(10):
        kept = [grid[x][y] for y in range(ROWS) if y not in full]
        out[x] = [0] * len(full) + kept
    return out, len(full)


def random_board(rng, fill_from=8):
    grid = [[0] * ROWS for _ in range(10)]
    for x in range(10):
        top = int(rng.integers(fill_from, ROWS + 1))
        for y in range(top, ROWS):
            grid[x][y] = int(rng.random() < 0.75)
    return grid


def test_fit_place_clear_match_brute_force():
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(60):
        grid = random_board(rng)
        cols = cols_of(grid)
        for t in range(7):
            for (o, x) in O.lut(t):
                b = O.Board(cols)
                legal, drop = b.fit(t, o, x)
                bd = brute_drop(grid, t, o, x)
                assert legal == (bd is not None)
                if not legal:
                    continue
                assert drop == bd
                g2 = [row[:] for row in grid]
                for (cx, cy) in piece_cells(t, o, x, drop):
                    g2[cx][cy] = 1
                g3, n = brute_clear(g2)
                b.place(t, o, x, drop)
                assert b.clear_lines() == n
                assert list(b.cols) == list(cols_of(g3))
                checked += 1
    assert checked > 1000


# --- tetris_engine_test.cpp --------------------------------------------------------------------------
def drain_one(e):  # tetris_engine_test.cpp:14-22
    t = e.queue[0]
    b = O.engine_board(e)
    for x in range(-2, 12):
        if b.available(t, TOP, x, 0):
            O.engine_place(e, TOP, x)
            break
    return t


def test_engine_plays():  # :24-32
    e = O.engine(123)
    assert not e.game_over
    assert O.engine_place(e, TOP, 3) == 0
    O.engine_hold(e)
    assert not e.game_over


def test_bag_is_a_permutation():  # :34-45 (and for many seeds / bags)
    e = O.engine(7)
    assert sorted(drain_one(e) for _ in range(7)) == list(range(7))
    for seed in range(50):
        e = O.engine(seed)
        seq = list(e.queue)
        # keep drawing through hold-free drains of an emptied board: the supply is independent of moves
        for _ in range(30):
            e2 = O.Engine.from_buffer_copy(e)
            O.engine_hold(e2)  # empty hold: advances the queue by one draw
            seq.append(e2.queue[5])
            e = e2
            e.held = 7
        for k in range(0, 28, 7):
            assert sorted(seq[k:k + 7]) == list(range(7))


def test_engine_reset_and_determinism():  # :49-74
    e = O.engine(1)
    assert not e.game_over and e.held == 7 and len(e.queue) == 6
    a, b = O.engine(42), O.engine(42)
    assert [drain_one(a) for _ in range(20)] == [drain_one(b) for _ in range(20)]
    a, b = O.engine(1), O.engine(2)
    assert [drain_one(a) for _ in range(14)] != [drain_one(b) for _ in range(14)]


def test_engine_advance_and_hold():  # :76-96
    e = O.engine(3)
    nxt = e.queue[1]
    assert O.engine_place(e, TOP, 3) == 0
    assert e.queue[0] == nxt
    e = O.engine(8)
    first = e.queue[0]
    assert e.held == 7
    O.engine_hold(e)
    assert e.held == first
    swapped = e.queue[0]
    O.engine_hold(e)
    assert e.held == swapped and e.queue[0] == first


def test_engine_tops_out():  # :98-108
    e = O.engine(5)
    guard = 0
    while not e.game_over and guard < 1000:
        O.engine_place(e, TOP, 3)
        guard += 1
    assert e.game_over and guard < 1000


def test_bag_draw_order_closed_form():
    # tetris_engine.hpp:116-129: identity fill, Fisher-Yates i = 6..1 with bounded(i+1) = next() % (i+1),
    # then the bag is consumed from the back. Restated over the (pinned) rng stream.
    for seed in (0, 1, 99, 2**63 + 5):
        st = [0, 0, 0, 0]
        import ctypes as C
        s = (C.c_uint64 * 4)()
        O.lib().ta3o_rng_seed(seed, s)
        stream = O.rng_outputs_from_state(list(s), 40)
        it = iter(stream)
        pieces = []
        for _ in range(2):
            bag = list(range(7))
            for i in range(6, 0, -1):
                j = next(it) % (i + 1)
                bag[i], bag[j] = bag[j], bag[i]
            pieces += bag[::-1]
        e = O.engine(seed)
        assert list(e.queue) == pieces[:6]
        del st
