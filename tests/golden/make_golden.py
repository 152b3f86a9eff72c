"""Writes normals_golden.json: the oracle's fp32 normals for fixed (seed, path)
pairs, as float32 bit patterns.  Run from the repository root after building
the oracle:  python tests/golden/make_golden.py"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

SEEDS = [0, 67, 0x123456789ABCDEF0, 2**64 - 1]
PATHS = [0, 1, 2**32 - 1, 2**32, 2**63 + 5]
N = 13

cases = []
for s in SEEDS:
    for p in PATHS:
        z = oracle.normals_f32(s, p, 1, N)[0]
        cases.append({"seed": hex(s), "path": hex(p), "n": N,
                      "bits": [f"0x{int(b):08x}" for b in z.view("uint32")]})
out = {"stream": "Philox4x32-10 key=(seed lo, seed hi), counter=(block, 0, path lo, path hi); "
                 "Box-Muller pairs (x0,x1), (x2,x3) with the fp32 polynomial kernels",
       "cases": cases}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "normals_golden.json"), "w") as f:
    json.dump(out, f, indent=1)
print(f"wrote {len(cases)} cases")
