"""The oracle against the committed golden vectors (tests/golden), which
need neither the reference tree nor a GPU.

tvl1_ref.npz holds flows computed by the reference's own compiled
lib/tvl1flow (tests/golden/make_tvl1_golden.py); the oracle's TV-L1
restatement must reproduce them bit for bit."""
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def tvl1_cases():
    z = np.load(os.path.join(GOLDEN, "tvl1_ref.npz"))
    n = len({k.split("_")[0] for k in z.files})
    for i in range(n):
        tau, lam, theta, ns, fs, zf, wp, eps = z[f"c{i}_prm"]
        kw = dict(tau=float(tau), lam=float(lam), theta=float(theta), nscales=int(ns), fscale=int(fs),
                  zfactor=float(zf), warps=int(wp), eps=float(eps))
        yield i, z[f"c{i}_I0"], z[f"c{i}_I1"], z[f"c{i}_u"], z[f"c{i}_v"], kw


@pytest.mark.parametrize("case", list(tvl1_cases()), ids=lambda c: f"case{c[0]}")
def test_tvl1_oracle_reproduces_reference_vectors(case):
    _, I0, I1, u, v, kw = case
    uo, vo, _ = O.tvl1(I0, I1, **kw)
    assert np.array_equal(uo, u) and np.array_equal(vo, v)
    assert np.abs(u).max() > 0.01      # a flow, not a zero field
