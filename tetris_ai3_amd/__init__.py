"""MI355X (gfx950) beam-search evaluator: Python binding of the C-ABI in include/ta3h.h.

Mirrors the reference's evaluator interface (ta3_gpu/incl/ta3/gpu/model_eval.hpp):
  ModelEvaluator.eval(weights, num_models, seeds, max_moves) -> fitness, model-major (:46-51)
  ModelEvaluator.eval(weights, seeds, max_moves)             -> single-model overload (:54-60)
  simulate(evaluator, weights, num_models, games, seed, max_moves) (:84-114)
All work runs in libta3h.so on the GPU; there is no CPU fallback. Loading fails loudly when the
library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libta3h.so")

MODEL_V11 = 0
MODEL_F16 = 1
MODEL_F32W = 3
PHASES_AUTO = 0xFFFFFFFF  # set_phases(None): first-phase budget from max_moves (the default)

_lib = None


class TrainConfig(C.Structure):
    _fields_ = [("kind", C.c_int), ("islands", C.c_uint32), ("population", C.c_uint32), ("games", C.c_uint32),
                ("max_moves", C.c_uint32), ("generations", C.c_uint32), ("sigma0", C.c_double),
                ("init_std", C.c_double), ("seed", C.c_uint64)]


class IterationReport(C.Structure):
    _fields_ = [("iteration", C.c_uint64), ("best_fitness", C.c_double), ("last_mean_fitness", C.c_double),
                ("champion_fitness", C.c_double), ("device_ms", C.c_double), ("iteration_seed", C.c_uint64)]


EVAL_CALLBACK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_float), C.c_uint32, C.c_uint32,
                            C.POINTER(C.c_uint64), C.c_uint32, C.c_uint32, C.POINTER(C.c_float))


class HipError(RuntimeError):
    """A failure reported by the evaluator library ("hip: <what>: <message>" or an argument error)."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        f32p = C.POINTER(C.c_float)
        u32p = C.POINTER(C.c_uint32)
        u64p = C.POINTER(C.c_uint64)
        sig = {
            "ta3h_create": (C.c_int, [C.POINTER(vp)]),
            "ta3h_destroy": (None, [vp]),
            "ta3h_last_error": (C.c_char_p, [vp]),
            "ta3h_num_params": (C.c_int, [C.c_int]),
            "ta3h_set_model": (C.c_int, [vp, C.c_int]),
            "ta3h_set_launch": (C.c_int, [vp, C.c_int, C.c_int, C.c_uint32]),
            "ta3h_set_beam_widths": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32]),
            "ta3h_eval": (C.c_int, [vp, f32p, C.c_uint32, C.c_uint32, u64p, C.c_uint32, C.c_uint32, f32p]),
            "ta3h_stage": (C.c_int, [vp, f32p, C.c_uint32, C.c_uint32, u64p, C.c_uint32, C.c_uint32]),
            "ta3h_launch": (C.c_int, [vp]),
            "ta3h_sync": (C.c_int, [vp]),
            "ta3h_fetch": (C.c_int, [vp, f32p, u32p]),
            "ta3h_last_kernel_ms": (C.c_int, [vp, f32p]),
            "ta3h_fetch_counters": (C.c_int, [vp, u64p]),
            "ta3h_set_timeline": (C.c_int, [vp, C.c_int]),
            "ta3h_set_move_log": (C.c_int, [vp, C.c_int]),
            "ta3h_fetch_moves": (C.c_int, [vp, C.POINTER(C.c_uint8)]),
            "ta3h_fetch_timeline": (C.c_int, [vp, u64p, C.POINTER(C.c_double)]),
            "ta3h_launch_info": (C.c_int, [vp, u32p, u32p, u32p]),
            "ta3h_device_bits_canary": (C.c_int, [C.POINTER(C.c_int)]),
            "ta3h_feature_norms": (C.c_int, [f32p, C.c_uint32]),
            "ta3h_set_garbage": (C.c_int, [vp, C.c_int]),
            "ta3h_set_phases": (C.c_int, [vp, C.c_uint32]),
            "ta3h_model_values": (C.c_int, [C.c_int, f32p, u32p, u32p, C.POINTER(C.c_uint8), C.c_uint32, f32p]),
            "ta3h_graph_prepare": (C.c_int, [vp, C.c_uint32, C.c_uint32, u64p, C.c_uint32, C.c_uint32]),
            "ta3h_graph_run": (C.c_int, [vp, f32p, f32p, u32p]),
            "ta3h_trainer_create": (C.c_int, [C.POINTER(TrainConfig), C.POINTER(vp)]),
            "ta3h_trainer_destroy": (None, [vp]),
            "ta3h_trainer_last_error": (C.c_char_p, [vp]),
            "ta3h_trainer_run_iteration": (C.c_int, [vp, vp, EVAL_CALLBACK, vp, C.POINTER(IterationReport)]),
            "ta3h_trainer_champion": (C.c_int, [vp, C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_double)]),
            "ta3h_trainer_save": (C.c_int, [vp, C.c_char_p]),
            "ta3h_trainer_load": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
            "ta3h_trainer_island_state": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "ta3h_champ_write": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.c_uint64]),
            "ta3h_champ_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_uint64)]),
            "ta3h_pruner_num_params": (C.c_int, []),
            "ta3h_pruner_quant_info": (C.c_int, [f32p, C.c_float, C.POINTER(C.c_int)]),
            "ta3h_set_pruner": (C.c_int, [vp, f32p, C.c_float, C.c_uint32, C.c_uint32]),
            "ta3h_set_trace": (C.c_int, [vp, C.c_uint64, C.c_uint32]),
            "ta3h_fetch_trace": (C.c_int, [vp, u32p, C.c_uint64, u64p, u64p]),
            "ta3h_pruner_grad": (C.c_int, [vp, f32p, u32p, C.c_uint32, C.c_float, f32p, f32p]),
            "ta3h_pruner_train": (C.c_int, [vp, f32p, u32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_float,
                                            C.c_float, C.c_uint64, f32p]),
            "ta3h_pruner_hidden": (C.c_int, [vp, f32p, u32p, C.c_uint32, f32p]),
            "ta3h_pruner_logits": (C.c_int, [vp, f32p, C.c_float, u32p, C.c_uint32, f32p]),
            "ta3h_eval_noalloc": (C.c_int, [C.c_int, vp, C.c_uint32, vp, C.c_uint32, C.c_uint32, vp, vp, vp]),
            "ta3h_replay": (C.c_int, [C.c_uint64, C.c_int, C.POINTER(C.c_uint8), C.c_uint32, u32p,
                                      C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]),
            "ta3h_tables": (None, [C.POINTER(C.c_int8), C.POINTER(C.c_uint8), C.POINTER(C.c_uint8),
                                   C.POINTER(C.c_uint8)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def num_params(kind: int = MODEL_V11) -> int:
    return lib().ta3h_num_params(kind)


class ModelEvaluator:
    """One evaluation context (a HIP stream + grow-only device buffers); not thread-safe, like the
    reference's ModelEvaluator (model_eval.hpp:21-26)."""

    def __init__(self, model: int = MODEL_V11):
        self._h = C.c_void_p()
        if lib().ta3h_create(C.byref(self._h)) != 0:
            raise HipError("ta3h_create failed")
        self.model = model
        self._items, self._max_moves = 0, 0
        self._garbage, self._pruned, self._traced, self._logged = False, False, False, False
        self._widths = (16, 32, 32)
        self._check(lib().ta3h_set_model(self._h, model))
        self.set_launch(DEFAULT_WAVES_PER_GAME.get(model, 2), True, 0)

    def close(self):
        if self._h:
            lib().ta3h_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            raise HipError(lib().ta3h_last_error(self._h).decode())

    @property
    def num_params(self) -> int:
        return num_params(self.model)

    def set_phases(self, first_phase_moves: int | None):
        """two-phase schedule: every game to first_phase_moves, then only the survivors, the games of the models
        with the most survivors first (0 = one phase; None = max_moves / 32 clamped to [64, 512], the default).
        Results never change."""
        self._check(lib().ta3h_set_phases(self._h, PHASES_AUTO if first_phase_moves is None else int(first_phase_moves)))

    def set_garbage(self, enabled: bool = True):
        """start every game of the following launches with 0-4 seeded garbage rows"""
        self._check(lib().ta3h_set_garbage(self._h, int(enabled)))
        self._garbage = bool(enabled)

    def set_beam_widths(self, widths=(16, 32, 32)):
        """beam widths per level for the following launches (at most 16 / 32 / 32, the reference's)"""
        k0, k1, k2 = (int(v) for v in widths)
        self._check(lib().ta3h_set_beam_widths(self._h, k0, k1, k2))
        self._widths = (k0, k1, k2)

    def set_launch(self, waves_per_game: int = 4, persistent: bool = True, grid_blocks: int = 0):
        """waves cooperating on one game (1, 2 or 4), persistent work queue or one workgroup per item, persistent
        grid size in workgroups (0 = occupancy limit). Results never depend on these."""
        self._check(lib().ta3h_set_launch(self._h, waves_per_game, int(persistent), grid_blocks))

    def _args(self, weights, num_models, seeds):
        w = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
        s = np.ascontiguousarray(seeds, dtype=np.uint64).reshape(-1)
        if num_models is None:
            num_models = 1
        P = self.num_params
        if w.size != num_models * P:
            raise HipError(f"eval: weights.size() must equal num_models * NUM_PARAMS ({num_models} * {P})")
        return w, int(num_models), s

    def eval(self, weights, num_models_or_seeds, seeds=None, max_moves=None):
        """eval(weights, num_models, seeds, max_moves) or eval(weights, seeds, max_moves)."""
        if max_moves is None:  # single-model overload (model_eval.hpp:54-60)
            num_models, seeds, max_moves = 1, num_models_or_seeds, seeds
        else:
            num_models = num_models_or_seeds
        w, m, s = self._args(weights, num_models, seeds)
        out = np.zeros(m * s.size, dtype=np.float32)
        self._items, self._max_moves = m * s.size, int(max_moves)
        self._check(lib().ta3h_eval(self._h, _ptr(w, C.c_float), m, self.num_params, _ptr(s, C.c_uint64), s.size,
                                    int(max_moves), _ptr(out, C.c_float)))
        return out

    # --- device steps, for callers that keep inputs resident and time the device work ---------------
    def stage(self, weights, num_models, seeds, max_moves):
        w, m, s = self._args(weights, num_models, seeds)
        self._items = m * s.size
        self._max_moves = int(max_moves)
        self._check(lib().ta3h_stage(self._h, _ptr(w, C.c_float), m, self.num_params, _ptr(s, C.c_uint64), s.size,
                                     int(max_moves)))

    def launch(self):
        self._check(lib().ta3h_launch(self._h))

    def sync(self):
        self._check(lib().ta3h_sync(self._h))

    def fetch(self):
        fit = np.zeros(self._items, dtype=np.float32)
        pl = np.zeros(self._items, dtype=np.uint32)
        self._check(lib().ta3h_fetch(self._h, _ptr(fit, C.c_float), _ptr(pl, C.c_uint32)))
        return fit, pl

    def fetch_counters(self):
        """per-item (candidates fitted, candidates legal, candidate lanes issued) of the last launch"""
        out = np.zeros(3 * self._items, dtype=np.uint64)
        self._check(lib().ta3h_fetch_counters(self._h, _ptr(out, C.c_uint64)))
        return out.reshape(-1, 3)

    def set_timeline(self, enabled: bool = True):
        self._check(lib().ta3h_set_timeline(self._h, int(enabled)))

    def fetch_timeline(self):
        """per-item (start, end) in seconds since the first start of the last launch"""
        out = np.zeros(2 * self._items, dtype=np.uint64)
        rate = C.c_double(0)
        self._check(lib().ta3h_fetch_timeline(self._h, _ptr(out, C.c_uint64), C.byref(rate)))
        t = out.reshape(-1, 2).astype(np.float64)
        return (t - t[:, 0].min()) / rate.value

    def set_move_log(self, enabled: bool = True):
        """record every committed move of the following launches (fetch_moves)"""
        self._check(lib().ta3h_set_move_log(self._h, int(enabled)))
        self._logged = bool(enabled)

    def fetch_moves(self):
        """move log of the last launch: uint8 [items, max_moves], byte i of an item = its move i as
        hold << 7 | rotation << 5 | (x + 2), for i below the item's placement count (the rest undefined)"""
        out = np.zeros(self._items * self._max_moves, dtype=np.uint8)
        self._check(lib().ta3h_fetch_moves(self._h, _ptr(out, C.c_uint8)))
        return out.reshape(self._items, self._max_moves)

    def last_kernel_ms(self) -> float:
        v = C.c_float(0)
        self._check(lib().ta3h_last_kernel_ms(self._h, C.byref(v)))
        return float(v.value)

    def launch_info(self):
        g, w, c = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
        self._check(lib().ta3h_launch_info(self._h, C.byref(g), C.byref(w), C.byref(c)))
        return {"grid_blocks": g.value, "waves_per_game": w.value, "num_cus": c.value}

    # --- one generation as a captured HIP graph ----------------------------------------------------
    def graph_prepare(self, num_models: int, seeds, max_moves: int):
        """captures weights-in, search, fitness-out for this (num_models, seeds, max_moves)"""
        s = np.ascontiguousarray(seeds, dtype=np.uint64).reshape(-1)
        self._graph_items = int(num_models) * s.size
        self._graph_models = int(num_models)
        self._items = self._graph_items
        self._max_moves = int(max_moves)
        self._check(lib().ta3h_graph_prepare(self._h, int(num_models), self.num_params, _ptr(s, C.c_uint64), s.size,
                                             int(max_moves)))

    def graph_run(self, weights):
        """one replay with new weights: (fitness, placements)"""
        w = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
        if w.size != self._graph_models * self.num_params:
            raise HipError("graph_run: weights must hold num_models * NUM_PARAMS values")
        fit = np.zeros(self._graph_items, dtype=np.float32)
        pl = np.zeros(self._graph_items, dtype=np.uint32)
        self._check(lib().ta3h_graph_run(self._h, _ptr(w, C.c_float), _ptr(fit, C.c_float), _ptr(pl, C.c_uint32)))
        return fit, pl

    # --- learned pruner ------------------------------------------------------------------------------
    def set_pruner(self, pruner_w, act_range: float, k: int = 8, ply_mask: int = 0b1100):
        """pruned beam for the following launches (ply_mask bit 1, 2: middle levels, bit 3: leaves);
        pruner_w None restores the full beam"""
        if pruner_w is None:
            self._check(lib().ta3h_set_pruner(self._h, None, 1.0, 34, 0))
            self._pruned = False
            return
        w = _pruner_w(pruner_w)
        self._check(lib().ta3h_set_pruner(self._h, _ptr(w, C.c_float), float(act_range), int(k), int(ply_mask)))
        self._pruned, self._traced = True, False

    def set_trace(self, cap_records: int, every: int = 1):
        """full-beam traces of every `every`-th move for the following launches (cap_records 0 = off)"""
        self._check(lib().ta3h_set_trace(self._h, int(cap_records), int(every)))
        self._trace_cap = int(cap_records)
        self._traced = int(cap_records) > 0 and int(every) > 0
        if self._traced:
            self._pruned = False

    def fetch_trace(self, cap_records: int | None = None):
        """(records uint32 [n, 11], produced) of the last launch; n <= the capacity set by set_trace"""
        produced, copied = C.c_uint64(0), C.c_uint64(0)
        self._check(lib().ta3h_fetch_trace(self._h, None, 0, C.byref(produced), C.byref(copied)))
        n = min(produced.value, getattr(self, "_trace_cap", 0))
        if cap_records is not None:
            n = min(n, int(cap_records))
        out = np.zeros((n, 11), dtype=np.uint32)
        if n:
            self._check(lib().ta3h_fetch_trace(self._h, _ptr(out, C.c_uint32), n, C.byref(produced), C.byref(copied)))
        return out[:copied.value], produced.value

    def pruner_grad(self, pruner_w, samples, pos_weight: float = 1.0):
        """(batch-mean loss, gradient float32[21249]) over samples [n, 11] on the device"""
        w = _pruner_w(pruner_w)
        smp = _samples(samples)
        g = np.zeros(pruner_num_params(), dtype=np.float32)
        loss = C.c_float(0)
        self._check(lib().ta3h_pruner_grad(self._h, _ptr(w, C.c_float), _ptr(smp, C.c_uint32), smp.shape[0],
                                           float(pos_weight), _ptr(g, C.c_float), C.byref(loss)))
        return float(loss.value), g

    def pruner_train(self, pruner_w, samples=None, n: int | None = None, epochs: int = 1, batch: int = 512,
                     lr: float = 1e-3, pos_weight: float = 1.0, seed: int = 0):
        """Adam on the device; samples None trains on the last launch's trace buffer (first n records).
        Returns (new weights, per-epoch mean loss)."""
        w = _pruner_w(pruner_w).copy()
        if samples is None:
            smp_p = None
            if n is None:
                raise HipError("pruner_train: n is required when training on the trace buffer")
        else:
            smp = _samples(samples)
            n = smp.shape[0] if n is None else min(int(n), smp.shape[0])
            smp_p = _ptr(smp, C.c_uint32)
        losses = np.zeros(max(int(epochs), 1), dtype=np.float32)
        self._check(lib().ta3h_pruner_train(self._h, _ptr(w, C.c_float), smp_p, int(n), int(epochs), int(batch),
                                            float(lr), float(pos_weight), int(seed), _ptr(losses, C.c_float)))
        return w, losses[:int(epochs)]

    def pruner_hidden(self, pruner_w, samples):
        w = _pruner_w(pruner_w)
        smp = _samples(samples)
        out = np.zeros((smp.shape[0], 64), dtype=np.float32)
        self._check(lib().ta3h_pruner_hidden(self._h, _ptr(w, C.c_float), _ptr(smp, C.c_uint32), smp.shape[0],
                                             _ptr(out, C.c_float)))
        return out

    def pruner_logits(self, pruner_w, act_range: float, samples):
        """inference logits as the pruned search computes them (integer grids on BF16 MFMA)"""
        w = _pruner_w(pruner_w)
        smp = _samples(samples)
        out = np.zeros(smp.shape[0], dtype=np.float32)
        self._check(lib().ta3h_pruner_logits(self._h, _ptr(w, C.c_float), float(act_range), _ptr(smp, C.c_uint32),
                                             smp.shape[0], _ptr(out, C.c_float)))
        return out


class Trainer:
    """CMA-ES island loop around the evaluator (ta3_trainer's Trainer2 shape): one batched generation of every
    island per graph replay. evaluate(weights[M, P], seeds, max_moves) -> fitness[M * G] replaces the device
    evaluator when given (e.g. the CPU restatement, to run the identical loop on the host)."""

    def __init__(self, kind=MODEL_V11, islands=4, population=50, games=50, max_moves=2000, generations=10,
                 sigma0=0.3, init_std=0.5, seed=0):
        self.cfg = TrainConfig(kind, islands, population, games, max_moves, generations, sigma0, init_std, seed)
        self._h = C.c_void_p()
        if lib().ta3h_trainer_create(C.byref(self.cfg), C.byref(self._h)) != 0:
            raise HipError("trainer: invalid configuration")
        self._set_shape(kind, islands, population, games, max_moves, generations)

    def _set_shape(self, kind, islands, population, games, max_moves, generations):
        self.kind, self.num_params = kind, num_params(kind)
        self.cfg_islands, self.cfg_population, self.cfg_games, self.cfg_generations = islands, population, games, generations
        self.cfg_max_moves = max_moves

    def close(self):
        if self._h:
            lib().ta3h_trainer_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_iteration(self, evaluator: "ModelEvaluator | None" = None, evaluate=None) -> dict:
        rep = IterationReport()
        cb = EVAL_CALLBACK()
        if evaluate is not None:
            def _cb(ctx, w, m, p, seeds, games, max_moves, out):
                try:
                    wa = np.ctypeslib.as_array(w, shape=(m * p,)).copy()
                    sa = np.ctypeslib.as_array(seeds, shape=(games,)).copy()
                    fit = np.asarray(evaluate(wa, m, sa, max_moves), dtype=np.float32).reshape(-1)
                    np.ctypeslib.as_array(out, shape=(m * games,))[:] = fit
                    return 0
                except Exception:
                    return 1
            cb = EVAL_CALLBACK(_cb)
        h = evaluator._h if evaluator is not None else None
        rc = lib().ta3h_trainer_run_iteration(self._h, h, cb, None, C.byref(rep))
        if rc != 0:
            raise HipError(lib().ta3h_trainer_last_error(self._h).decode())
        return {f: getattr(rep, f) for f, _ in IterationReport._fields_}

    def champion(self):
        out = np.zeros(self.num_params, dtype=np.float64)
        fit = C.c_double(0)
        if lib().ta3h_trainer_champion(self._h, _ptr(out, C.c_double), out.size, C.byref(fit)) != 0:
            raise HipError("trainer: no champion yet")
        return out, fit.value

    def island_state(self, island: int):
        mean = np.zeros(self.num_params, dtype=np.float64)
        sigma = C.c_double(0)
        if lib().ta3h_trainer_island_state(self._h, island, _ptr(mean, C.c_double), C.byref(sigma)) != 0:
            raise HipError("trainer: no such island")
        return mean, sigma.value

    def save(self, path: str) -> None:
        """checkpoint between iterations (config, seeds, champion, every island's CMA-ES state)"""
        if lib().ta3h_trainer_save(self._h, os.fsencode(path)) != 0:
            raise HipError(f"trainer: cannot write {path}")

    @classmethod
    def load(cls, path: str) -> "Trainer":
        """a trainer that continues exactly where the checkpoint was saved"""
        h = C.c_void_p()
        if lib().ta3h_trainer_load(os.fsencode(path), C.byref(h)) != 0:
            raise HipError(f"trainer: cannot load {path}")
        t = cls.__new__(cls)
        t._h = h
        t.cfg = None
        with open(path, "rb") as f:  # the config follows the magic and the version: i32 kind, then 5 x u32
            head = f.read(36)
        kind, islands, population, games, max_moves, generations = np.frombuffer(head[12:36], dtype="<i4")
        t._set_shape(int(kind), int(islands), int(population), int(games), int(max_moves), int(generations))
        return t


def champ_write(path: str, w) -> None:
    """.champ: u64 count + count little-endian doubles (trainer2.cpp:251-265)"""
    a = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
    if lib().ta3h_champ_write(os.fsencode(path), _ptr(a, C.c_double), a.size) != 0:
        raise HipError(f"champ_write: cannot write {path}")


def champ_read(path: str) -> np.ndarray:
    n = C.c_uint64(0)
    if lib().ta3h_champ_read(os.fsencode(path), None, 0, C.byref(n)) != 0:
        raise HipError(f"champ_read: cannot read {path}")
    out = np.zeros(n.value, dtype=np.float64)
    if lib().ta3h_champ_read(os.fsencode(path), _ptr(out, C.c_double), out.size, C.byref(n)) != 0:
        raise HipError(f"champ_read: malformed {path}")
    return out


def pruner_num_params() -> int:
    return lib().ta3h_pruner_num_params()


def _pruner_w(w):
    a = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
    if a.size != pruner_num_params():
        raise HipError(f"pruner weights must hold {pruner_num_params()} values")
    return a


def _samples(samples):
    smp = np.ascontiguousarray(samples, dtype=np.uint32).reshape(-1, 11)
    if smp.shape[0] == 0:
        raise HipError("no samples")
    return smp


def pruner_quant_info(pruner_w, act_range: float):
    """grid exponents (e1, e2, e3, d) of the integer inference form"""
    w = _pruner_w(pruner_w)
    info = (C.c_int * 4)()
    if lib().ta3h_pruner_quant_info(_ptr(w, C.c_float), float(act_range), info) != 0:
        raise HipError("pruner_quant_info: invalid arguments")
    return tuple(info)


def pruner_init(seed: int = 0) -> np.ndarray:
    """He-initialised master weights (zero biases)"""
    rng = np.random.default_rng(seed)
    w = np.zeros(pruner_num_params(), dtype=np.float32)
    o = 0
    for fan_in, fan_out in ((265, 64), (64, 64), (64, 1)):
        w[o:o + fan_in * fan_out] = rng.normal(0.0, np.sqrt(2.0 / fan_in), fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    return w


# launch shapes measured fastest on MI355X per network (full CMA-ES generation); autotune() refines them per batch
DEFAULT_WAVES_PER_GAME = {MODEL_V11: 2, MODEL_F16: 4, MODEL_F32W: 4}  # V11 at 5 waves/SIMD, F16 at 4

# tuned launch shapes: (network, models, games, max_moves, garbage, kernel mode) -> (waves per game, grid, phase)
_TUNED: dict = {}


def _tune_key(evaluator: "ModelEvaluator", num_models: int, games: int, max_moves: int) -> str:
    mode = "pruned" if evaluator._pruned else ("trace" if evaluator._traced or evaluator._logged else "plain")
    widths = "-".join(str(k) for k in evaluator._widths)
    return (f"{evaluator.model}:{int(num_models)}x{int(games)}:{int(max_moves)}:g{int(evaluator._garbage)}:{mode}"
            f":w{widths}")


def _cache_path() -> str:
    return os.environ.get("TA3H_AUTOTUNE_CACHE", "")


def _load_cache() -> dict:
    path = _cache_path()
    if path and os.path.exists(path):
        import json
        with open(path) as f:
            return {k: tuple(v) for k, v in json.load(f).items()}
    return {}


def autotune(evaluator: "ModelEvaluator", weights, num_models: int, seeds, max_moves: int,
             candidates=None, phases=(None, 32, 128), repeats: int = 2, use_cache: bool = True):
    """Picks the launch shape (waves per game, grid) and first-phase budget for this batch shape and keeps it on
    `evaluator` (none of them changes a result). The choice is cached per (network, models x games, max_moves,
    garbage rows, kernel mode) in this process and, when TA3H_AUTOTUNE_CACHE names a JSON file, across processes;
    a cached shape is applied without timing. Otherwise every candidate is timed `repeats` times on the batch
    (device time of the search) and the fastest kept. The pruned and trace kernels always run 4 waves per game, so
    only their phase budget is tuned. Leaves the batch staged. Returns {"key", "chosen", "cached", "timings_ms"}."""
    key = _tune_key(evaluator, num_models, len(seeds), max_moves)
    if use_cache:
        if key not in _TUNED:
            _TUNED.update({k: v for k, v in _load_cache().items() if k not in _TUNED})
        if key in _TUNED:
            wpg, grid, phase = _TUNED[key]
            evaluator.set_launch(wpg, True, grid)
            evaluator.set_phases(phase)
            evaluator.stage(weights, num_models, seeds, max_moves)
            return {"key": key, "chosen": list(_TUNED[key]), "cached": True, "timings_ms": {}}
    if candidates is None: # (waves per game, grid)
        candidates = ((1, 0), (2, 0), (4, 0))
    if evaluator._pruned or evaluator._traced or evaluator._logged:
        candidates = ((4, 0),)
    evaluator.stage(weights, num_models, seeds, max_moves)
    timings = {}
    for wpg, grid in candidates:
        for phase in phases:
            evaluator.set_launch(wpg, True, grid)
            evaluator.set_phases(phase)
            best = float("inf")
            for _ in range(max(1, int(repeats))):
                evaluator.launch()
                evaluator.sync()
                best = min(best, evaluator.last_kernel_ms())
            timings[(wpg, grid, phase)] = best
    chosen = min(timings, key=timings.get)
    evaluator.set_launch(chosen[0], True, chosen[1])
    evaluator.set_phases(chosen[2])
    _TUNED[key] = chosen
    path = _cache_path()
    if use_cache and path:
        import json
        merged = _load_cache()
        merged[key] = chosen
        with open(path, "w") as f:
            json.dump({k: list(v) for k, v in merged.items()}, f, indent=1, sort_keys=True)
    return {"key": key, "chosen": list(chosen), "cached": False,
            "timings_ms": {f"wpg{k[0]}_grid{k[1]}_phase{'auto' if k[2] is None else k[2]}": v
                           for k, v in timings.items()}}


E_SCRATCH = 4
NOALLOC_SCRATCH_BYTES = 256 << 20


def eval_noalloc(kind: int, d_weights: int, num_models: int, d_seeds: int, games: int, max_moves: int,
                 d_fitness: int, d_placements: int = 0, stream: int = 0) -> int:
    """ta3h_eval_noalloc on device addresses (integers, e.g. torch's data_ptr()) and a HIP stream handle (0 = the
    null stream): enqueues the evaluation without allocating or synchronising; returns the status code (0 = ok,
    E_SCRATCH = the batch exceeds the static scratch array)."""
    return lib().ta3h_eval_noalloc(int(kind), C.c_void_p(d_weights), int(num_models), C.c_void_p(d_seeds), int(games),
                                   int(max_moves), C.c_void_p(d_fitness), C.c_void_p(d_placements or None),
                                   C.c_void_p(stream or None))


def replay(seed: int, moves, garbage: bool = False):
    """host replay of one game's move log (bytes hold << 7 | rotation << 5 | (x + 2)): (boards uint32 [n, 10] after
    each move, pieces placed [n], rows cleared [n]); n < len(moves) after a top-out"""
    m = np.ascontiguousarray(moves, dtype=np.uint8).reshape(-1)
    boards = np.zeros((max(m.size, 1), 10), dtype=np.uint32)
    pieces = np.zeros(max(m.size, 1), dtype=np.uint8)
    cleared = np.zeros(max(m.size, 1), dtype=np.uint8)
    n = lib().ta3h_replay(int(seed) & ((1 << 64) - 1), int(garbage), _ptr(m, C.c_uint8), m.size, _ptr(boards, C.c_uint32),
                          _ptr(pieces, C.c_uint8), _ptr(cleared, C.c_uint8))
    if n < 0:
        raise HipError("replay: a move does not fit")
    return boards[:n], pieces[:n], cleared[:n]


def expand_seeds(seed: int, games: int) -> np.ndarray:
    """gpu::simulate's seed expansion: seeds[g] = xoshiro256**{seed}() in order (model_eval.hpp:92-95)."""
    mask = (1 << 64) - 1
    s = seed & mask
    st = []
    for _ in range(4):  # splitmix64 seeding (xoshiro256ss.hpp:20-28)
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        st.append(z ^ (z >> 31))
    out = []
    for _ in range(games):
        r = (((st[1] * 5) & mask) << 7 | ((st[1] * 5) & mask) >> 57) & mask
        out.append((r * 9) & mask)
        t = (st[1] << 17) & mask
        st[2] ^= st[0]
        st[3] ^= st[1]
        st[1] ^= st[2]
        st[0] ^= st[3]
        st[2] ^= t
        st[3] = ((st[3] << 45) | (st[3] >> 19)) & mask
    return np.array(out, dtype=np.uint64)


def simulate(evaluator: ModelEvaluator, model_weights, num_models: int, games: int, seed: int, max_moves: int):
    """Batched simulate (model_eval.hpp:84-103): num_models models over `games` shared games seeded from `seed`."""
    return evaluator.eval(model_weights, num_models, expand_seeds(seed, games), max_moves)


def model_values(kind: int, weights, boards, clears, held) -> np.ndarray:
    """leaf values of boards [n, 10] as the search scores candidate boards (one model, host layout)"""
    w = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
    if w.size != num_params(kind):
        raise HipError(f"model_values: {num_params(kind)} weights expected")
    b = np.ascontiguousarray(boards, dtype=np.uint32).reshape(-1, 10)
    c = np.ascontiguousarray(clears, dtype=np.uint32).reshape(-1)
    h = np.ascontiguousarray(held, dtype=np.uint8).reshape(-1)
    out = np.zeros(b.shape[0], dtype=np.float32)
    if lib().ta3h_model_values(kind, _ptr(w, C.c_float), _ptr(b, C.c_uint32), _ptr(c, C.c_uint32), _ptr(h, C.c_uint8),
                               b.shape[0], _ptr(out, C.c_float)) != 0:
        raise HipError("model_values failed")
    return out


def device_bits_canary() -> bool:
    ok = C.c_int(0)
    if lib().ta3h_device_bits_canary(C.byref(ok)) != 0:
        raise HipError("device_bits_canary: HIP error")
    return bool(ok.value)


def feature_norms(n: int) -> np.ndarray:
    """device values [i, (sqrt(i)/100, i/20, i/200, i/5, i/40, i/100)] for i < n"""
    out = np.zeros(6 * n, dtype=np.float32)
    if lib().ta3h_feature_norms(_ptr(out, C.c_float), n) != 0:
        raise HipError("feature_norms: HIP error")
    return out.reshape(n, 6)


def tables():
    """Host copy of the compile-time tables: ({piece: [(o, x), ...]}, trees[2][4] lists of node dicts)."""
    ox = np.zeros(162 * 2, dtype=np.int8)
    off = np.zeros(8, dtype=np.uint8)
    tree = np.zeros(2 * 4 * 16 * 6, dtype=np.uint8)
    cnt = np.zeros(8, dtype=np.uint8)
    lib().ta3h_tables(_ptr(ox, C.c_int8), _ptr(off, C.c_uint8), _ptr(tree, C.c_uint8), _ptr(cnt, C.c_uint8))
    lut = {p: [(int(ox[2 * i]), int(ox[2 * i + 1])) for i in range(off[p], off[p + 1])] for p in range(7)}
    trees = []
    for h in range(2):
        layers = []
        for l in range(4):
            layer = []
            for n in range(int(cnt[h * 4 + l])):
                v = tree[((h * 4 + l) * 16 + n) * 6:((h * 4 + l) * 16 + n) * 6 + 6]
                layer.append(dict(slot=int(v[0]), parent=int(v[1]), child_begin=int(v[2]), child_end=int(v[3]),
                                  root_hold=bool(v[4]), held_slot=int(v[5])))
            layers.append(layer)
        trees.append(layers)
    return lut, trees
