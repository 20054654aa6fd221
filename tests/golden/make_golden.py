"""Writes the golden fixtures under tests/golden/ (run in the build container).

- v11_champion.json: the reference's shipped v11 champion weights, read from its data file
  models/v11_16_32_32_score_296.champ (format: u64 count, then count little-endian f64;
  ta3_trainer/src/trainer2.cpp:251-265). The file name records the champion's mean score_v1
  (296) under the reference's beam widths 16/32/32 -- an end-to-end known answer.
- parity_cases.json: explicit weight vectors (numpy PCG64, uniform(-2, 2) like the reference
  parity test ta3_gpu/tests/src/parity_test.cpp:20-32, but stored because std distributions
  differ across libraries), seeds and caps, with the oracle's fitness for each case.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"


def champion():
    path = os.path.join(REF, "models", "v11_16_32_32_score_296.champ")
    count = int(np.fromfile(path, dtype="<u8", count=1)[0])
    w = np.fromfile(path, dtype="<f8", offset=8)
    assert count == 97 and w.shape == (97,)
    with open(os.path.join(HERE, "v11_champion.json"), "w") as f:
        json.dump({"source": "models/v11_16_32_32_score_296.champ", "count": count,
                   "weights_f64": [float(v) for v in w]}, f, indent=0)


def parity_cases():
    sys.path.insert(0, ROOT)
    import oracle as O
    cases = []
    rng = np.random.Generator(np.random.PCG64(20261203))
    seeds = [1, 2, 3, 4]  # parity_test.cpp:36-39
    for kind, P in ((O.MODEL_V11, 97), (O.MODEL_F16, 1633)):
        for ms in range(3):
            w = rng.uniform(-2.0, 2.0, size=P).astype(np.float32)
            if kind == O.MODEL_F16:
                w = (rng.normal(0.0, 0.5, size=P)).astype(np.float32)
            fit = O.eval_host(kind, w, seeds, 200)
            cases.append({"kind": kind, "weights": [float(v) for v in w], "seeds": seeds, "max_moves": 200,
                          "fitness": [float(v) for v in fit]})
    with open(os.path.join(HERE, "parity_cases.json"), "w") as f:
        json.dump({"cases": cases}, f)


if __name__ == "__main__":
    champion()
    parity_cases()
