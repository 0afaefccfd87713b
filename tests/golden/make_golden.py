"""Writes the golden vectors under tests/golden/ from the CPU oracle.

rng.json: HybridTaus streams from fixed states and the per-thread initial
states of the three seeding formulas (capture fw, capture bw, map eval) for a
few thread indices.  The CPU tests check that the oracle keeps producing them
(tests/test_oracle.py); the GPU parity tests compare the HIP kernels' streams
with the oracle's, not with this file (tests/test_gpu_parity.py exposes the
first uniform of every thread through a homogeneous collision capture).

Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

streams = []
for st in ([12345, 67890, 13579, 24680], [129, 130, 131, 132], [1 << 29, 7, 1 << 30 - 1, 999]):
    u, _ = O.rng_uniforms(st, 16)
    streams.append({"state": st, "uniforms": [float(v) for v in u]})
seed_states = []
for mode in (0, 1, 2):
    for seeds, first in (([12345, 67890, 13579, 24680], 0), ([129, 129, 129, 129], 1000000)):
        s = O.seed_states(mode, seeds, first, 8)
        seed_states.append({"mode": mode, "seeds": seeds, "first": first, "states": s.tolist()})
with open(os.path.join(HERE, "rng.json"), "w") as f:
    json.dump({"streams": streams, "seed_states": seed_states}, f, indent=1)
print("wrote rng.json")
