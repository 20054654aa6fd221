"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a scalar C++ restatement of the reference's CPU evaluation path
(ta3_gpu/src/common/model_eval_host.cpp:26-54). Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module; the product package
(tetris_ai3_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libta3oracle.so")

MODEL_V11, MODEL_F16, MODEL_HEIGHT, MODEL_F32W = 0, 1, 2, 3
NO_PIECE = 7
REF_WIDTHS = (16, 32, 32)


class Engine(C.Structure):
    _fields_ = [
        ("cols", C.c_uint32 * 10),
        ("queue", C.c_uint32 * 6),
        ("held", C.c_uint32),
        ("game_over", C.c_uint32),
        ("bag", C.c_uint32 * 7),
        ("bag_pos", C.c_uint32),
        ("rng", C.c_uint64 * 4),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("moves", C.c_uint64),
        ("applies", C.c_uint64),
        ("legal", C.c_uint64),
        ("reapplies", C.c_uint64),
        ("placements", C.c_uint64),
        ("dead_mid", C.c_uint64),
        ("dead_leaf", C.c_uint64),
        ("fallback_mid_best", C.c_uint64),
        ("fallback_leaf_best", C.c_uint64),
        ("repeat_applies", C.c_uint64),  # of applies / legal: in repeated lists of a middle level or the leaves
        ("repeat_legal", C.c_uint64),
        ("reuse_applies", C.c_uint64),  # of applies / legal: roots the previous move scored (no line cleared)
        ("reuse_legal", C.c_uint64),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def build() -> None:
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32p = C.POINTER(C.c_uint32)
        f32p = C.POINTER(C.c_float)
        u64p = C.POINTER(C.c_uint64)
        u8p = C.POINTER(C.c_uint8)
        ip = C.POINTER(C.c_int)
        sig = {
            "ta3o_num_params": (C.c_int, [C.c_int]),
            "ta3o_num_features": (C.c_int, [C.c_int]),
            "ta3o_rng_seed": (None, [C.c_uint64, u64p]),
            "ta3o_rng_next": (C.c_uint64, [u64p]),
            "ta3o_expand_seeds": (None, [C.c_uint64, C.c_uint32, u64p]),
            "ta3o_board_available": (C.c_int, [u32p, C.c_int, C.c_int, C.c_int, C.c_int]),
            "ta3o_board_drop_location": (C.c_int, [u32p, C.c_int, C.c_int, C.c_int, C.c_int]),
            "ta3o_board_place": (None, [u32p, C.c_int, C.c_int, C.c_int, C.c_int]),
            "ta3o_board_clear_lines": (C.c_int, [u32p]),
            "ta3o_board_full_lines": (C.c_int, [u32p]),
            "ta3o_board_holes": (C.c_int, [u32p]),
            "ta3o_board_fit": (C.c_int, [u32p, C.c_int, C.c_int, C.c_int, ip]),
            "ta3o_piece_shape": (None, [C.c_int, C.c_int, u32p, ip]),
            "ta3o_engine_reset": (None, [C.POINTER(Engine), C.c_uint64]),
            "ta3o_engine_place": (C.c_int, [C.POINTER(Engine), C.c_int, C.c_int]),
            "ta3o_engine_hold": (None, [C.POINTER(Engine)]),
            "ta3o_lut_count": (C.c_int, [C.c_int]),
            "ta3o_lut_entry": (None, [C.c_int, C.c_int, ip, ip]),
            "ta3o_plans": (C.c_int, [C.c_int, C.c_int, u8p, u8p, u8p]),
            "ta3o_tree": (None, [C.c_int, C.c_int, u8p, u8p]),
            "ta3o_ordered_bits": (C.c_uint32, [C.c_float]),
            "ta3o_ordered_bits_inv": (C.c_float, [C.c_uint32]),
            "ta3o_features": (None, [C.c_int, C.c_uint32, u32p, C.c_int, f32p]),
            "ta3o_model_value": (C.c_float, [C.c_int, f32p, C.c_uint32, u32p, C.c_int]),
            "ta3o_f32w_value_fp32": (C.c_float, [f32p, C.c_uint32, u32p, C.c_int]),
            "ta3o_set_garbage": (None, [C.c_int]),
            "ta3o_set_widths": (None, [u32p]),
            "ta3o_search_beam": (C.c_int, [C.POINTER(Engine), C.c_int, f32p, u32p, C.c_int,
                                           ip, ip, ip, f32p, C.POINTER(Counters)]),
            "ta3o_search_beam_pruned": (C.c_int, [C.POINTER(Engine), C.c_int, f32p, f32p, C.c_float, C.c_int, C.c_uint,
                                                  ip, ip, ip, f32p]),
            "ta3o_candidate_ops": (None, [C.c_int, f32p, u32p, C.c_int, u64p, C.POINTER(C.c_int32), u32p, f32p,
                                          f32p]),
            "ta3o_select_top": (C.c_uint32, [u32p, C.c_uint32, C.c_uint32, u32p, u32p]),
            "ta3o_set_select_linear": (None, [C.c_int]),
            "ta3o_search_exhaustive": (C.c_int, [C.POINTER(Engine), C.c_int, f32p, C.c_int, ip, ip, ip, f32p]),
            "ta3o_eval_host": (None, [C.c_int, f32p, u64p, C.c_uint32, C.c_uint32, f32p, C.POINTER(Counters)]),
            "ta3o_eval_host_mt": (None, [C.c_int, f32p, C.c_uint32, u64p, C.c_uint32, C.c_uint32, C.c_int,
                                         f32p, u32p, C.POINTER(Counters)]),
            "ta3o_play_trace": (C.c_int, [C.c_int, f32p, C.c_uint64, C.c_uint32, u8p, u32p, f32p]),
            "ta3o_check_deep_reuse": (C.c_int, [C.c_int, f32p, C.c_uint64, C.c_uint32, u64p]),
            "ta3o_pruner_num_params": (C.c_int, []),
            "ta3o_pruner_score": (C.c_float, [f32p, C.c_float, u32p, C.c_int, C.c_int, C.c_int, C.c_int]),
            "ta3o_pruner_scores": (None, [f32p, C.c_float, u32p, C.c_int, f32p]),
            "ta3o_pruner_quant_info": (None, [f32p, C.c_float, C.POINTER(C.c_int)]),
            "ta3o_pruner_loss_grad": (C.c_double, [f32p, u32p, C.c_int, C.c_double, C.c_int,
                                                   C.POINTER(C.c_double)]),
            "ta3o_eval_host_pruned_mt": (None, [C.c_int, f32p, C.c_uint32, u64p, C.c_uint32, C.c_uint32, C.c_int,
                                                f32p, C.c_float, C.c_int, C.c_uint, f32p, u32p,
                                                C.POINTER(Counters)]),
            "ta3o_trace_records": (C.c_uint64, [C.c_int, f32p, C.c_uint64, C.c_uint32, C.c_uint32, u32p,
                                                C.c_uint64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def num_params(kind: int) -> int:
    return lib().ta3o_num_params(kind)


# --- rng / seeds -------------------------------------------------------------------------------
def rng_outputs(seed: int, n: int) -> list[int]:
    st = (C.c_uint64 * 4)()
    lib().ta3o_rng_seed(seed, st)
    return [lib().ta3o_rng_next(st) for _ in range(n)]


def rng_outputs_from_state(state, n: int) -> list[int]:
    st = (C.c_uint64 * 4)(*state)
    return [lib().ta3o_rng_next(st) for _ in range(n)]


def expand_seeds(seed: int, games: int) -> np.ndarray:
    out = np.zeros(games, dtype=np.uint64)
    lib().ta3o_expand_seeds(seed, games, _p(out, C.c_uint64))
    return out


# --- board ---------------------------------------------------------------------------------------
class Board:
    """Board2 restatement handle: 10 u32 columns, bit y = row y (y=0 top)."""

    def __init__(self, cols=None):
        self.cols = np.zeros(10, dtype=np.uint32) if cols is None else np.array(cols, dtype=np.uint32)

    def _c(self):
        return _p(self.cols, C.c_uint32)

    def available(self, t, o, x, y):
        return bool(lib().ta3o_board_available(self._c(), t, o, x, y))

    def drop_location(self, t, o, x, y):
        return lib().ta3o_board_drop_location(self._c(), t, o, x, y)

    def place(self, t, o, x, y):
        lib().ta3o_board_place(self._c(), t, o, x, y)

    def drop_place(self, t, o, x, y):
        self.place(t, o, x, self.drop_location(t, o, x, y))

    def clear_lines(self):
        return lib().ta3o_board_clear_lines(self._c())

    def full_lines(self):
        return lib().ta3o_board_full_lines(self._c())

    def holes(self):
        return lib().ta3o_board_holes(self._c())

    def fit(self, t, o, x):
        d = C.c_int(0)
        legal = lib().ta3o_board_fit(self._c(), t, o, x, C.byref(d))
        return bool(legal), d.value


def piece_shape(t, o):
    cols = C.c_uint32(0)
    left = C.c_int(0)
    lib().ta3o_piece_shape(t, o, C.byref(cols), C.byref(left))
    return cols.value, left.value


# --- engine --------------------------------------------------------------------------------------
def engine(seed: int) -> Engine:
    e = Engine()
    lib().ta3o_engine_reset(C.byref(e), seed)
    return e


def engine_place(e: Engine, o: int, x: int) -> int:
    return lib().ta3o_engine_place(C.byref(e), o, x)


def engine_hold(e: Engine) -> None:
    lib().ta3o_engine_hold(C.byref(e))


def engine_board(e: Engine) -> Board:
    return Board(list(e.cols))


# --- tables --------------------------------------------------------------------------------------
def lut(piece: int):
    n = lib().ta3o_lut_count(piece)
    out = []
    for i in range(n):
        o = C.c_int(0)
        x = C.c_int(0)
        lib().ta3o_lut_entry(piece, i, C.byref(o), C.byref(x))
        out.append((o.value, x.value))
    return out


def plans(depth: int, held_empty: bool):
    n = 1 << depth
    slots = np.zeros(n * depth, dtype=np.uint8)
    held = np.zeros(n * depth, dtype=np.uint8)
    root = np.zeros(n, dtype=np.uint8)
    cnt = lib().ta3o_plans(depth, int(held_empty), _p(slots, C.c_uint8), _p(held, C.c_uint8), _p(root, C.c_uint8))
    return [(tuple(slots[p * depth:(p + 1) * depth]), tuple(held[p * depth:(p + 1) * depth]), bool(root[p]))
            for p in range(cnt)]


def tree(depth: int, held_empty: bool):
    n = 1 << depth
    nodes = np.zeros(depth * n * 6, dtype=np.uint8)
    counts = np.zeros(depth, dtype=np.uint8)
    lib().ta3o_tree(depth, int(held_empty), _p(nodes, C.c_uint8), _p(counts, C.c_uint8))
    layers = []
    for l in range(depth):
        layer = []
        for i in range(int(counts[l])):
            v = nodes[(l * n + i) * 6:(l * n + i) * 6 + 6]
            layer.append(dict(slot=int(v[0]), parent=int(v[1]), child_begin=int(v[2]), child_end=int(v[3]),
                              root_hold=bool(v[4]), held_slot=int(v[5])))
        layers.append(layer)
    return layers


def ordered_bits(v: float) -> int:
    return lib().ta3o_ordered_bits(v)


def ordered_bits_inv(s: int) -> float:
    return lib().ta3o_ordered_bits_inv(s)


# --- model ---------------------------------------------------------------------------------------
def features(kind: int, clears: int, cols, held_is_i: bool) -> np.ndarray:
    n = lib().ta3o_num_features(kind)
    out = np.zeros(n, dtype=np.float32)
    c = np.array(cols, dtype=np.uint32)
    lib().ta3o_features(kind, clears, _p(c, C.c_uint32), int(held_is_i), _p(out, C.c_float))
    return out


def f32w_value_fp32(w, clears: int, cols, held_is_i: bool) -> float:
    """the wide network in plain FP32 (no bfloat16 rounding)"""
    w = _f32(w)
    c = np.array(cols, dtype=np.uint32)
    return lib().ta3o_f32w_value_fp32(_p(w, C.c_float), clears, _p(c, C.c_uint32), int(held_is_i))


class garbage:
    """context manager: every game the oracle starts inside it begins with 0-4 garbage rows"""

    def __enter__(self):
        lib().ta3o_set_garbage(1)
        return self

    def __exit__(self, *exc):
        lib().ta3o_set_garbage(0)
        return False


class widths:
    """context manager: the games the oracle plays inside it search with these beam widths (k0, k1, k2)"""

    def __init__(self, ks):
        self.ks = np.array(ks, dtype=np.uint32)

    def __enter__(self):
        lib().ta3o_set_widths(_p(self.ks, C.c_uint32))
        return self

    def __exit__(self, *exc):
        lib().ta3o_set_widths(None)
        return False


def model_value(kind: int, w, clears: int, cols, held_is_i: bool) -> float:
    w = _f32(w if len(w) else np.zeros(1))
    c = np.array(cols, dtype=np.uint32)
    return lib().ta3o_model_value(kind, _p(w, C.c_float), clears, _p(c, C.c_uint32), int(held_is_i))


# --- search --------------------------------------------------------------------------------------
def search_beam(e: Engine, kind: int, w, widths=REF_WIDTHS, counters: Counters | None = None):
    w = _f32(w if len(w) else np.zeros(1))
    ks = np.array(widths, dtype=np.uint32)
    h, o, x = C.c_int(0), C.c_int(0), C.c_int(0)
    v = C.c_float(0)
    cnt = counters if counters is not None else Counters()
    r = lib().ta3o_search_beam(C.byref(e), kind, _p(w, C.c_float), _p(ks, C.c_uint32), len(ks),
                               C.byref(h), C.byref(o), C.byref(x), C.byref(v), C.byref(cnt))
    if r < 0:
        raise ValueError("bad beam widths")
    if r == 0:
        return None
    return dict(hold=bool(h.value), o=o.value, x=x.value, value=v.value)


def search_beam_pruned(e: Engine, kind: int, w, pruner_w, act_range: float, k: int, ply_mask: int):
    """one pruned search move at the reference widths (None when there is no move)"""
    w = _f32(w if len(w) else np.zeros(1))
    pw = _f32(pruner_w)
    h, o, x = C.c_int(0), C.c_int(0), C.c_int(0)
    v = C.c_float(0)
    r = lib().ta3o_search_beam_pruned(C.byref(e), kind, _p(w, C.c_float), _p(pw, C.c_float), float(act_range), int(k),
                                      int(ply_mask), C.byref(h), C.byref(o), C.byref(x), C.byref(v))
    if r == 0:
        return None
    return dict(hold=bool(h.value), o=o.value, x=x.value, value=v.value)


def select_top(scores, k: int):
    """select_top (beam.hpp:244-265): (slots, scores, scores after the call with the selected ones zeroed)"""
    sc = np.array(scores, dtype=np.uint32)
    slots = np.zeros(max(k, 1), dtype=np.uint32)
    top = np.zeros(max(k, 1), dtype=np.uint32)
    n = lib().ta3o_select_top(_p(sc, C.c_uint32), sc.size, k, _p(slots, C.c_uint32), _p(top, C.c_uint32))
    return slots[:n].tolist(), top[:n].tolist(), sc


class select_linear:
    """context manager: select_top runs the repeated-argmax loop for every beam width (else widths above 64
    take an equivalent single sort)"""

    def __enter__(self):
        lib().ta3o_set_select_linear(1)
        return self

    def __exit__(self, *exc):
        lib().ta3o_set_select_linear(0)
        return False


def search_exhaustive(e: Engine, kind: int, w, depth: int):
    w = _f32(w if len(w) else np.zeros(1))
    h, o, x = C.c_int(0), C.c_int(0), C.c_int(0)
    v = C.c_float(0)
    r = lib().ta3o_search_exhaustive(C.byref(e), kind, _p(w, C.c_float), depth, C.byref(h), C.byref(o),
                                     C.byref(x), C.byref(v))
    if r < 0:
        raise ValueError("bad depth")
    if r == 0:
        return None
    return dict(hold=bool(h.value), o=o.value, x=x.value, value=v.value)


# --- evaluation ----------------------------------------------------------------------------------
def eval_host(kind: int, w, seeds, max_moves: int, counters: Counters | None = None) -> np.ndarray:
    """eval_host for ONE model: fitness per game (model_eval_host.cpp:26-54)."""
    w = _f32(w if len(w) else np.zeros(1))
    s = np.ascontiguousarray(seeds, dtype=np.uint64)
    out = np.zeros(len(s), dtype=np.float32)
    cnt = counters if counters is not None else Counters()
    lib().ta3o_eval_host(kind, _p(w, C.c_float), _p(s, C.c_uint64), len(s), max_moves, _p(out, C.c_float),
                         C.byref(cnt))
    return out


def eval_host_mt(kind: int, w, num_models: int, seeds, max_moves: int, threads: int = 0,
                 counters: Counters | None = None):
    """Batched CPU evaluation over (model, game) items on host threads; returns (fitness, placements)."""
    w = _f32(w if len(w) else np.zeros(1))
    s = np.ascontiguousarray(seeds, dtype=np.uint64)
    items = num_models * len(s)
    out = np.zeros(items, dtype=np.float32)
    pl = np.zeros(items, dtype=np.uint32)
    cnt = counters if counters is not None else Counters()
    if threads <= 0:
        threads = os.cpu_count() or 1
    lib().ta3o_eval_host_mt(kind, _p(w, C.c_float), num_models, _p(s, C.c_uint64), len(s), max_moves, threads,
                            _p(out, C.c_float), _p(pl, C.c_uint32), C.byref(cnt))
    return out, pl


def check_deep_reuse(kind: int, w, seed: int, max_moves: int) -> dict:
    """one game with the content check of the lists of levels 1 and 2 the search counts as taken from the previous
    move (ta3_oracle.cpp DeepCheck): lists counted, those whose content no list of the previous search one level
    down had, moves, reuse_applies"""
    w = _f32(w if len(w) else np.zeros(1))
    out = np.zeros(4, dtype=np.uint64)
    lib().ta3o_check_deep_reuse(kind, _p(w, C.c_float), seed, max_moves, _p(out, C.c_uint64))
    return {"counted": int(out[0]), "mismatched": int(out[1]), "moves": int(out[2]), "reuse_applies": int(out[3])}


def play_trace(kind: int, w, seed: int, max_moves: int):
    """One game with every move recorded: (moves[n,4] = hold,o,x+2,cleared(255=DIED), boards[n,10], fitness)."""
    w = _f32(w if len(w) else np.zeros(1))
    moves = np.zeros(max_moves * 4, dtype=np.uint8)
    boards = np.zeros(max_moves * 10, dtype=np.uint32)
    fit = C.c_float(0)
    n = lib().ta3o_play_trace(kind, _p(w, C.c_float), seed, max_moves, _p(moves, C.c_uint8),
                              _p(boards, C.c_uint32), C.byref(fit))
    return moves[:n * 4].reshape(n, 4), boards[:n * 10].reshape(n, 10), fit.value


# --- placement-proposal network (pruner) ----------------------------------------------------------
def pruner_num_params() -> int:
    return lib().ta3o_pruner_num_params()


def pruner_score(pw, act_range, cols, piece, o, x, ply) -> float:
    """integer-grid inference logit of one candidate"""
    pw = _f32(pw)
    c = np.array(cols, dtype=np.uint32)
    return lib().ta3o_pruner_score(_p(pw, C.c_float), float(act_range), _p(c, C.c_uint32), piece, o, x, ply)


def pruner_scores(pw, act_range, samples):
    """integer-grid inference logits of trace records [n, 11]"""
    pw = _f32(pw)
    smp = np.ascontiguousarray(samples, dtype=np.uint32).reshape(-1, 11)
    out = np.zeros(smp.shape[0], dtype=np.float32)
    lib().ta3o_pruner_scores(_p(pw, C.c_float), float(act_range), _p(smp, C.c_uint32), smp.shape[0],
                             _p(out, C.c_float))
    return out


def pruner_quant_info(pw, act_range):
    """grid exponents (e1, e2, e3, d) of the quantized network"""
    pw = _f32(pw)
    info = (C.c_int * 4)()
    lib().ta3o_pruner_quant_info(_p(pw, C.c_float), float(act_range), info)
    return tuple(info)


def pruner_loss_grad(pw, samples, pos_weight=1.0, rounded=True, with_grad=True):
    """(loss, grad float64[P]) of the batch-mean weighted BCE over trace records [n, 11] uint32"""
    pw = _f32(pw)
    smp = np.ascontiguousarray(samples, dtype=np.uint32).reshape(-1, 11)
    g = np.zeros(pruner_num_params(), dtype=np.float64) if with_grad else None
    loss = lib().ta3o_pruner_loss_grad(_p(pw, C.c_float), _p(smp, C.c_uint32), smp.shape[0], float(pos_weight),
                                       int(rounded), _p(g, C.c_double) if with_grad else None)
    return loss, g


def eval_host_pruned_mt(kind, w, num_models, seeds, max_moves, pruner_w, act_range, k, ply_mask, threads=0):
    w = _f32(w if len(w) else np.zeros(1))
    pw = _f32(pruner_w)
    s = np.ascontiguousarray(seeds, dtype=np.uint64)
    items = num_models * len(s)
    out = np.zeros(items, dtype=np.float32)
    pl = np.zeros(items, dtype=np.uint32)
    cnt = Counters()
    if threads <= 0:
        threads = os.cpu_count() or 1
    lib().ta3o_eval_host_pruned_mt(kind, _p(w, C.c_float), num_models, _p(s, C.c_uint64), len(s), max_moves,
                                   threads, _p(pw, C.c_float), float(act_range), k, ply_mask, _p(out, C.c_float),
                                   _p(pl, C.c_uint32),
                                   C.byref(cnt))
    return out, pl, cnt


def candidate_ops(kind, w, samples):
    """instrumented op counts of candidates given as trace records [n, 11]: (ops per part summed (fit, place +
    clear, features, network), drop or -1 per record, boards [n, 10], features [n, 16], values [n])"""
    w = _f32(w)
    smp = np.ascontiguousarray(samples, dtype=np.uint32).reshape(-1, 11)
    n = smp.shape[0]
    ops = np.zeros(4, dtype=np.uint64)
    legal = np.zeros(n, dtype=np.int32)
    boards = np.zeros((n, 10), dtype=np.uint32)
    feats = np.zeros((n, 16), dtype=np.float32)
    vals = np.zeros(n, dtype=np.float32)
    lib().ta3o_candidate_ops(kind, _p(w, C.c_float), _p(smp, C.c_uint32), n, _p(ops, C.c_uint64),
                             _p(legal, C.c_int32), _p(boards, C.c_uint32), _p(feats, C.c_float), _p(vals, C.c_float))
    return ops, legal, boards, feats, vals


def trace_records(kind, w, seed, max_moves, every):
    """all trace records of one full-beam game: uint32 [n, 11]"""
    w = _f32(w if len(w) else np.zeros(1))
    n = lib().ta3o_trace_records(kind, _p(w, C.c_float), seed, max_moves, every, None, 0)
    out = np.zeros(n * 11, dtype=np.uint32)
    lib().ta3o_trace_records(kind, _p(w, C.c_float), seed, max_moves, every, _p(out, C.c_uint32), n)
    return out.reshape(-1, 11)
