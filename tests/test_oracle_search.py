"""Pins the oracle's placement LUT, hold plans, plan trees, keys and beam search.

Known answers and invariants from the reference tests:
  ta3_ai/tests/src/search/search_test.cpp (LUT counts :60-67, naive hold replay :132-229),
  ta3_ai/tests/src/search/beam_test.cpp (tree structure :44-100, keys :103-107,
  full-width beam == exhaustive :113-141, determinism :144-169).
What the mathematics fixes beyond them:
  select_top (beam.hpp:244-265) == a stable sort by score descending, slot ascending, zeros dropped;
  the dead-end fallbacks (beam.hpp:439-440 per leaf list, :549-551 per kept node with no legal child) on
  near-top-out boards: at depth 2 the full-width beam equals the exhaustive search (search.hpp:117-150), and at
  depth 3 it equals a brute-force walk of the plan tree that scores every leaf and every dead end.
"""
import numpy as np
import pytest

import oracle as O

I, OP, T, S, Z, J, L = range(7)
HELD_NONE = 7


# --- placement LUT ---------------------------------------------------------------------------------
def test_lut_counts_pinned():  # search_test.cpp:60-67
    assert [len(O.lut(p)) for p in range(7)] == [17, 9, 34, 17, 17, 34, 34]


def test_lut_entries_in_range_and_ordered():  # search_test.cpp:40-56; placements.hpp:189-205 order
    for p in range(7):
        prev = (-1, -99)
        for (o, x) in O.lut(p):
            cols, left = O.piece_shape(p, o)
            n = (cols.bit_length() + 3) // 4
            assert -left <= x <= 10 - left - n
            assert (o, x) > prev  # orientation enum, then x ascending
            prev = (o, x)


def test_lut_dedup_loses_no_resting_board():
    """Brute force: the deduped LUT reaches exactly the set of resting boards that all 4 orientations x
    all in-range window offsets reach (placements.hpp:31-37), on random boards away from the top."""
    rng = np.random.default_rng(11)
    for _ in range(40):
        cols = np.zeros(10, dtype=np.uint32)
        for x in range(10):
            h = int(rng.integers(0, 12))
            cols[x] = ((1 << 23) - 1) & ~((1 << (23 - h)) - 1)
            cols[x] &= ~np.uint32(1 << int(rng.integers(12, 23)))
        for p in range(7):
            def results(pairs):
                out = set()
                for (o, x) in pairs:
                    b = O.Board(cols)
                    legal, d = b.fit(p, o, x)
                    if legal:
                        b.place(p, o, x, d)
                        b.clear_lines()
                        out.add(tuple(b.cols))
                return out
            full = []
            for o in range(4):
                c, left = O.piece_shape(p, o)
                n = (c.bit_length() + 3) // 4
                full += [(o, x) for x in range(-left, 10 - left - n + 1)]
            assert results(O.lut(p)) == results(full)


# --- hold plans: naive replay (search_test.cpp:132-229) ---------------------------------------------
def naive_plans(depth, held_empty):
    out = []
    for d in range(1 << depth):
        cur, hld, empty, k = 1, 0, held_empty, 1
        slots, held = [], []
        root_hold = False
        for i in range(depth):
            if d & (1 << i):
                if empty:
                    hld, cur, k, empty = cur, 1 + k, k + 1, False
                else:
                    cur, hld = hld, cur
                if i == 0:
                    root_hold = True
            slots.append(cur)
            held.append(HELD_NONE if empty else hld)
            if i + 1 < depth:
                cur, k = 1 + k, k + 1
        out.append((tuple(slots), tuple(held), root_hold))
    return out


@pytest.mark.parametrize("depth", [2, 3, 4, 5])
@pytest.mark.parametrize("held_empty", [False, True])
def test_plans_match_naive_replay(depth, held_empty):
    assert O.plans(depth, held_empty) == naive_plans(depth, held_empty)


def test_plan_invariants_depth4():  # search_test.cpp:183-229
    pe = O.plans(4, True)
    never = [p for p in pe if p[0] == (1, 2, 3, 4)]
    assert never and all(h == (7, 7, 7, 7) for (_, h, _) in never)
    pf = O.plans(4, False)
    never = [p for p in pf if p[0] == (1, 2, 3, 4)]
    assert never and all(h == (0, 0, 0, 0) for (_, h, _) in never)
    assert any(r and h[0] == 1 for (_, h, r) in pe)


# --- plan trees (beam_test.cpp:44-100) ---------------------------------------------------------------
@pytest.mark.parametrize("depth", [2, 3, 4])
@pytest.mark.parametrize("held_empty", [False, True])
def test_tree_consistent_and_covers_plans(depth, held_empty):
    layers = O.tree(depth, held_empty)
    for l in range(depth - 1):
        nxt = 0
        for n, node in enumerate(layers[l]):
            assert node["child_begin"] == nxt and node["child_end"] >= node["child_begin"]
            for c in range(node["child_begin"], node["child_end"]):
                assert layers[l + 1][c]["parent"] == n
            nxt = node["child_end"]
        assert nxt == len(layers[l + 1])
    plans = O.plans(depth, held_empty)
    assert len(layers[depth - 1]) == len(plans)
    for slots, held, root_hold in plans:
        begin, end = 0, len(layers[0])
        for l in range(depth):
            hit = [n for n in range(begin, end) if layers[l][n]["slot"] == slots[l]
                   and (l != 0 or layers[l][n]["root_hold"] == root_hold)]
            assert len(hit) >= 1
            node = layers[l][hit[0]]
            assert node["held_slot"] == held[l]
            begin, end = node["child_begin"], node["child_end"]


def test_tree_layer_counts_reference_depth():  # derived constants: layer counts {2,4,8,16}, caps 2/2/2
    for he in (False, True):
        layers = O.tree(4, he)
        assert [len(x) for x in layers] == [2, 4, 8, 16]
        assert max(n["child_end"] - n["child_begin"] for l in range(3) for n in layers[l]) == 2


# --- ordered keys (beam_test.cpp:103-107) ----------------------------------------------------------
def test_ordered_bits_monotone_and_invertible():
    assert O.ordered_bits(-1.0) < O.ordered_bits(-0.5) < O.ordered_bits(0.0) < O.ordered_bits(2.0)
    assert O.ordered_bits_inv(O.ordered_bits(3.25)) == 3.25
    assert O.ordered_bits_inv(O.ordered_bits(-7.5)) == -7.5
    vals = np.sort(np.random.default_rng(1).normal(0, 100, 2000).astype(np.float32))
    keys = [O.ordered_bits(float(v)) for v in vals]
    assert all(a <= b for a, b in zip(keys, keys[1:]))
    assert all(k != 0 for k in keys)


# --- beam search ------------------------------------------------------------------------------------
def play_compare(depth, seed, moves):
    """a beam wide enough to keep everything reproduces the exhaustive search's best value
    (beam_test.cpp:113-141; the reference runs it whenever DEPTH <= 3)."""
    root = 68
    widths = [root] + [root * 2 * 34 * (2 * 34) ** (i - 1) for i in range(1, depth - 1)]
    e = O.engine(seed)
    done = 0
    for _ in range(moves):
        if e.game_over:
            break
        ref = O.search_exhaustive(e, O.MODEL_HEIGHT, [], depth)
        beam = O.search_beam(e, O.MODEL_HEIGHT, [], widths)
        assert (ref is None) == (beam is None)
        if ref is None:
            break
        assert ref["value"] == beam["value"]
        if beam["hold"]:
            O.engine_hold(e)
        assert O.engine_place(e, beam["o"], beam["x"]) != -1
        done += 1
    return done


def test_full_width_beam_matches_exhaustive_depth2():
    assert play_compare(2, 99, 40) == 40


def test_full_width_beam_matches_exhaustive_depth3():
    assert play_compare(3, 99, 12) == 12


@pytest.mark.slow
def test_full_width_beam_matches_exhaustive_depth4():
    assert play_compare(4, 99, 2) == 2


def test_full_width_beam_matches_exhaustive_v11_depth3():
    """the same equality with the V11 network (mirror-invariant features, so canonical() is inert)"""
    rng = np.random.default_rng(3)
    w = rng.uniform(-2, 2, 97).astype(np.float32)
    e = O.engine(17)
    widths = [68, 68 * 68]
    for _ in range(6):
        ref = O.search_exhaustive(e, O.MODEL_V11, w, 3)
        beam = O.search_beam(e, O.MODEL_V11, w, widths)
        assert ref["value"] == beam["value"]
        if beam["hold"]:
            O.engine_hold(e)
        if O.engine_place(e, beam["o"], beam["x"]) == -1:
            break


def test_beam_plays_deterministic_games():  # beam_test.cpp:144-169
    def play(seed):
        e = O.engine(seed)
        trace = 0
        for _ in range(120):
            if e.game_over:
                break
            r = O.search_beam(e, O.MODEL_HEIGHT, [])
            if r is None:
                break
            if r["hold"]:
                O.engine_hold(e)
            if O.engine_place(e, r["o"], r["x"]) == -1:
                break
            trace = (trace * 31 + r["o"] * 7 + (r["x"] + 2)) & ((1 << 64) - 1)
        return trace
    assert play(3) == play(3)
    assert play(42) == play(42)
    assert play(3) != play(42)


def test_beam_move_is_legal_and_value_bounded_by_exhaustive():
    """the decoded root move fits the committed board (sanity of decode_move), and a narrow beam's value is at
    most the exhaustive search's at the same depth: every key the beam considers is a leaf of the full tree or a
    dead end that the exhaustive search also scores"""
    rng = np.random.default_rng(8)
    w = rng.uniform(-2, 2, 97).astype(np.float32)
    e = O.engine(4)
    for _ in range(30):
        r = O.search_beam(e, O.MODEL_V11, w)
        if r is None:
            break
        narrow = O.search_beam(e, O.MODEL_V11, w, (16, 32))
        assert narrow["value"] <= O.search_exhaustive(e, O.MODEL_V11, w, 3)["value"]
        piece = e.held if r["hold"] and e.held != 7 else (e.queue[1] if r["hold"] else e.queue[0])
        b = O.engine_board(e)
        legal, _ = b.fit(piece, r["o"], r["x"])
        assert legal
        if r["hold"]:
            O.engine_hold(e)
        if O.engine_place(e, r["o"], r["x"]) == -1:
            break


# --- select_top against a brute-force stable sort (beam.hpp:244-265) ----------------------------------
@pytest.mark.parametrize("linear", [False, True])
def test_select_top_matches_stable_sort_with_ties(linear):
    """the k best nonzero scores, highest first, the LOWEST slot first among equal scores; the selected scores
    are zeroed. Scores come from a handful of values so that ties are everywhere, at every width the search uses
    (16, 32) and wider than the argmax loop's range (65, 100)."""
    rng = np.random.default_rng(2024)
    for trial in range(300):
        n = int(rng.choice([1, 3, 34, 68, 100, 1088, 2176]))
        k = int(rng.choice([1, 2, 16, 32, 64, 65, 100]))
        scores = rng.choice(np.array([0, 0, 1, 7, 0x80000000, 0xBF800000, 0xFFFFFFFF], dtype=np.uint32), n)
        want = sorted((i for i in range(n) if scores[i] != 0), key=lambda i: (-int(scores[i]), i))[:k]
        if linear:
            with O.select_linear():
                slots, top, after = O.select_top(scores, k)
        else:
            slots, top, after = O.select_top(scores, k)
        assert slots == want, (trial, n, k)
        assert top == [int(scores[i]) for i in want]
        rest = scores.copy()
        rest[want] = 0
        assert np.array_equal(after, rest)


def test_full_width_argmax_path_matches_exhaustive():
    """the full-width equality above, with every select_top run as the repeated-argmax loop that the reference
    widths (<= 64) take"""
    with O.select_linear():
        assert play_compare(2, 7, 25) == 25
        assert play_compare(3, 7, 6) == 6


# --- dead ends: near-top-out boards ------------------------------------------------------------------
FIELD = (1 << 23) - 1


def near_top_out(seed, rng):
    """an engine of seed `seed` (hold used or not) whose board is filled to 1-3 rows below the top, one gap per
    row (no row is full), so that pieces soon find no room: kept nodes without a legal child and leaf lists
    without a legal placement occur"""
    e = O.engine(seed)
    if rng.integers(0, 2):
        O.engine_hold(e)
    cols = [FIELD & ~((1 << int(rng.integers(1, 4))) - 1) for _ in range(10)]
    for y in range(23):
        cols[(y * 7 + seed) % 10] &= ~(1 << y)
    for x in range(10):
        e.cols[x] = cols[x]
    e.game_over = 0 if O.Board(cols).available(e.queue[0], 0, 3, 0) else 1
    return e


def _hist_added(h, cleared):  # clear_hist_t::added (model_defs.hpp:53-55)
    return h if cleared == 0 else h + (1 << (8 * (cleared - 1)))


def plan_tree_walk(e, kind, w, depth):
    """Brute force over the whole plan tree (beam.hpp:69-126) at full width: the best of every leaf's value, of
    the depth-(depth-2) node's own value for each of its leaf lists with no legal placement (beam.hpp:439-440),
    and of a shallower node's own value when none of its child lists has a legal placement (beam.hpp:549-551).
    No selection, no keys, no orders: plain enumeration with the board primitives and the model."""
    slots = [e.held, e.queue[0]] + [e.queue[1 + i] for i in range(depth)]
    tree = O.tree(depth, e.held == 7)
    luts = {p: O.lut(p) for p in range(7)}
    best = []

    def held_i(n):
        return n["held_slot"] != 7 and slots[n["held_slot"]] == 0

    def children(board, clears, piece):
        for o, x in luts[piece]:
            b = O.Board(board)
            legal, d = b.fit(piece, o, x)
            if legal:
                b.place(piece, o, x, d)
                cleared = b.clear_lines()  # before the board is read
                yield list(b.cols), _hist_added(clears, cleared)

    def walk(level, node, board, clears):
        v = O.model_value(kind, w, clears, board, held_i(node))
        if level == depth - 1:
            best.append(v)
            return
        any_child = False
        for ci in range(node["child_begin"], node["child_end"]):
            child = tree[level + 1][ci]
            list_has = False
            for b2, c2 in children(board, clears, slots[child["slot"]]):
                list_has = any_child = True
                walk(level + 1, child, b2, c2)
            if level + 1 == depth - 1 and not list_has:
                best.append(v)
        if level + 1 < depth - 1 and not any_child:
            best.append(v)

    for root in tree[0]:
        for b2, c2 in children(list(e.cols), 0, slots[root["slot"]]):
            walk(0, root, b2, c2)
    return max(best) if best else None


def _dead_end_cases(n):
    rng = np.random.default_rng(77)
    w = rng.uniform(-2, 2, 97).astype(np.float32)
    for seed in range(n):
        e = near_top_out(seed, rng)
        if not e.game_over:
            yield e, O.MODEL_HEIGHT, []
            yield e, O.MODEL_V11, w


@pytest.mark.parametrize("linear", [False, True])
def test_leaf_dead_ends_full_width_depth2_equal_exhaustive(linear):
    """at depth 2 (roots, then leaves) the exhaustive search scores a root whose next piece does not fit exactly
    as the beam's leaf-list fallback does, so the values agree on boards where that decides the move"""
    total = O.Counters()
    cases = 0
    for e, kind, w in _dead_end_cases(60):
        ref = O.search_exhaustive(e, kind, w, 2)
        if linear:
            with O.select_linear():
                beam = O.search_beam(e, kind, w, (68,), total)
        else:
            beam = O.search_beam(e, kind, w, (68,), total)
        assert (ref is None) == (beam is None)
        if ref is not None:
            assert beam["value"] == ref["value"]
            cases += 1
    assert cases >= 40
    assert total.dead_leaf > 0 and total.fallback_leaf_best >= 5  # the fallback decided moves


@pytest.mark.parametrize("linear", [False, True])
def test_dead_ends_full_width_depth3_equal_plan_tree_walk(linear):
    """at depth 3 the full-width beam equals the brute-force plan-tree walk on near-top-out boards where both
    fallbacks (kept node without a legal child, leaf list without a legal placement) decide moves. The exhaustive
    search is no judge here: it scores a node whenever ONE of its plans dead-ends, the beam only when all of its
    child lists do (beam.hpp:549-551), and on some of these boards that changes the value."""
    total = O.Counters()
    cases = differ = 0
    for e, kind, w in _dead_end_cases(45):
        brute = plan_tree_walk(e, kind, w, 3)
        if linear:
            with O.select_linear():
                beam = O.search_beam(e, kind, w, (68, 68 * 68), total)
        else:
            beam = O.search_beam(e, kind, w, (68, 68 * 68), total)
        assert (brute is None) == (beam is None)
        if brute is None:
            continue
        assert beam["value"] == brute
        cases += 1
        ref = O.search_exhaustive(e, kind, w, 3)
        differ += ref["value"] != brute
    assert cases >= 30
    assert total.dead_mid > 0 and total.fallback_mid_best >= 2
    assert total.dead_leaf > 0 and total.fallback_leaf_best >= 2
    assert differ >= 1  # the reference's per-node rule is visible on these boards


def test_lists_taken_from_the_previous_move_have_its_content():
    """The 16-32-32-1 search counts level-1 and level-2 lists as taken from the previous move by structural
    identities (the committed root's kept level-1 node, its level-2 lists with the same piece and held-is-I; one
    level deeper likewise). Independently of those identities, every list counted must have the content (board,
    clear histogram, piece, held-is-I) of a list the previous search built one level down; a wrong child, rank
    or slot offset in the identities breaks this (checked by planting each in a copy of the oracle)."""
    rng = np.random.default_rng(11)
    kind = 1  # the 16-32-32-1 network (TA3O_MODEL_F16)
    counted = 0
    for k in range(3):
        w = rng.normal(0.0, 0.5, O.num_params(kind)).astype(np.float32)
        for seed in (5, 77):
            r = O.check_deep_reuse(kind, w, seed, 150)
            assert r["mismatched"] == 0, r
            counted += r["counted"]
    assert counted > 1000  # the games are long enough to take many lists
