"""ta3h_eval_noalloc (no counterpart in the reference): the allocation-free variant of ta3h_eval gives the same
fitness and placements as ta3h_eval, on two calls in a row (the second smaller, so stale scratch contents of the
first call would show), for every value network, on a caller's stream and on the null stream.

Device buffers and the stream come from the HIP runtime that libta3h.so itself loaded (through ctypes), so the
test never brings a second HIP runtime into the process."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import tetris_ai3_amd as T

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    """the libamdhip64 that libta3h.so is linked against, as mapped into this process"""
    T.lib()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    mine = [p for p in paths if "/torch/" not in p]
    assert mine, paths
    hip = C.CDLL(sorted(mine)[0])
    for name, args in {"hipMalloc": [C.POINTER(C.c_void_p), C.c_size_t], "hipFree": [C.c_void_p],
                       "hipMemcpy": [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int],
                       "hipStreamCreate": [C.POINTER(C.c_void_p)], "hipStreamDestroy": [C.c_void_p],
                       "hipStreamSynchronize": [C.c_void_p]}.items():
        getattr(hip, name).argtypes = args
        getattr(hip, name).restype = C.c_int
    return hip


class Dev:
    """a device copy of a host array (or an empty device array of that shape)"""

    def __init__(self, hip, host, copy=True):
        self.hip, self.host = hip, np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 4)) == 0
        if copy:
            assert hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def noalloc(hip, kind, w, M, seeds, cap, stream):
    dw = Dev(hip, np.asarray(w, dtype=np.float32))
    ds = Dev(hip, np.asarray(seeds, dtype=np.uint64))
    df = Dev(hip, np.full(M * len(seeds), np.nan, dtype=np.float32))
    dp = Dev(hip, np.zeros(M * len(seeds), dtype=np.uint32))
    rc = T.eval_noalloc(kind, dw.ptr.value, M, ds.ptr.value, len(seeds), cap, df.ptr.value, dp.ptr.value,
                        stream.value or 0)
    assert rc == 0
    assert hip.hipStreamSynchronize(stream) == 0
    fit, pl = df.get(), dp.get()
    for d in (dw, ds, df, dp):
        d.free()
    return fit, pl


@pytest.mark.parametrize("kind", [T.MODEL_V11, T.MODEL_F16, T.MODEL_F32W])
def test_noalloc_equals_eval_two_calls(kind):
    hip = hip_runtime()
    P = T.num_params(kind)
    rng = np.random.default_rng(90 + kind)
    scale = 0.5 if kind == T.MODEL_F32W else 1.0
    shapes = [(6, 5, 300), (3, 2, 150)] if kind != T.MODEL_F32W else [(4, 3, 200), (2, 2, 150)]
    ev = T.ModelEvaluator(model=kind)
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    for call, (M, G, cap) in enumerate(shapes):
        w = rng.normal(0.0, scale, M * P).astype(np.float32)
        seeds = T.expand_seeds(1000 + call, G)
        print(f"noalloc kind {kind} call {call}: {M} models x {G} games, cap {cap}", flush=True)
        want = ev.eval(w, M, seeds, cap)
        _, want_pl = ev.fetch()
        got, got_pl = noalloc(hip, kind, w, M, seeds, cap, stream)
        assert got.tobytes() == want.tobytes(), (call, got, want)
        assert np.array_equal(got_pl, want_pl)
        assert (got_pl >= 1).all() and (got_pl <= cap).all()
    hip.hipStreamDestroy(stream)
    ev.close()


def test_noalloc_null_stream_and_long_games():
    """the null stream, and games that outlive the first phase (parked and resumed from the static array)"""
    hip = hip_runtime()
    champ = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "v11_champion.json")))
    w = np.asarray(champ["weights_f64"], dtype=np.float32)
    seeds = T.expand_seeds(5, 3)
    ev = T.ModelEvaluator(model=T.MODEL_V11)
    want = ev.eval(w, 1, seeds, 400)
    _, want_pl = ev.fetch()
    ev.close()
    got, got_pl = noalloc(hip, T.MODEL_V11, w, 1, seeds, 400, C.c_void_p())
    assert got.tobytes() == want.tobytes()
    assert np.array_equal(got_pl, want_pl) and want_pl.max() > 128  # past the first phase's 128 moves
