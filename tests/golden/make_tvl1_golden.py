#!/usr/bin/env python3
"""Writes tests/golden/tvl1_ref.npz: TV-L1 flows computed by the reference's
own lib/tvl1flow (tvl1flow_lib.c, compiled from its sources into
oracle/_ref/libtvl1ref.so by oracle/Makefile) on small seeded frame pairs,
with one OpenMP thread (the stopping test's reduction order then is the
sequential one).

Run from the repository root after `make -C oracle` (needs the reference tree
to build oracle/_ref):   python3 tests/golden/make_tvl1_golden.py

Each case stores its two input frames, the parameters and the flow (u, v).
The frames are noisy frames of the procedural scene used throughout the
tests, so the fixtures carry no reference data other than these outputs.
"""
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

import oracle as O  # noqa: E402
from conftest import synth  # noqa: E402

# (w, h, seed, parameters): the flow parameters of the 960x540 videos
# (lambda 0.15, 10 warps, fscale 0), of tvl1flow-seq (lambda 0.2, 5 warps,
# fscale 1), and a short one with a finer stopping threshold
CASES = [
    (48, 40, 11, dict(tau=0.25, lam=0.15, theta=0.3, nscales=2, fscale=0, zfactor=0.5, warps=10, eps=0.01)),
    (64, 48, 12, dict(tau=0.25, lam=0.2, theta=0.3, nscales=3, fscale=1, zfactor=0.5, warps=5, eps=0.01)),
    (40, 36, 13, dict(tau=0.2, lam=0.1, theta=0.25, nscales=2, fscale=0, zfactor=0.6, warps=3, eps=0.005)),
]


def main():
    if O.ref_tvl1() is None:
        sys.exit("oracle/_ref/libtvl1ref.so is not built (make -C oracle with the reference tree present)")
    out = {}
    for i, (w, h, seed, kw) in enumerate(CASES):
        v = synth.add_noise(synth.video(w, h, 2, seed=seed), 20, seed + 100)
        I0, I1 = v[0, 0].copy(), v[1, 0].copy()
        u, vv = O.tvl1_reference(I0, I1, **kw)
        out[f"c{i}_I0"] = I0
        out[f"c{i}_I1"] = I1
        out[f"c{i}_u"] = u
        out[f"c{i}_v"] = vv
        out[f"c{i}_prm"] = np.array([kw["tau"], kw["lam"], kw["theta"], kw["nscales"], kw["fscale"],
                                     kw["zfactor"], kw["warps"], kw["eps"]], np.float64)
    np.savez_compressed(os.path.join(REPO, "tests", "golden", "tvl1_ref.npz"), **out)
    print("wrote", len(CASES), "cases")


if __name__ == "__main__":
    main()
