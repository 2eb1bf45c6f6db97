"""GPU TV-L1 against the CPU oracle restatement (itself bit-exact with the
reference's compiled lib/tvl1flow, see test_oracle_pins) and against the
reference library directly when oracle/_ref is present.  Per-pixel
arithmetic matches operation for operation; the stopping test sums the mean
squared update in a different order, so an iteration count may differ by one
where the error lands next to epsilon^2."""
import numpy as np
import pytest

import oracle as O
from conftest import bm4dof_amd as A, synth

pytestmark = pytest.mark.gpu


def _pair(w, h, seed):
    v = synth.add_noise(synth.video(w, h, 2, seed=seed), 20, seed + 1)
    return v[0, 0], v[1, 0]


@pytest.mark.parametrize("w,h,kw", [(64, 48, dict(nscales=3, warps=3)),
                                    (96, 72, dict(nscales=4, warps=5, lam=0.2)),
                                    (80, 60, dict(nscales=3, warps=2, fscale=1))])
def test_tvl1_matches_oracle(ctx, w, h, kw):
    I0, I1 = _pair(w, h, w + h)
    u, v, it = ctx.tvl1(I0, I1, **kw)
    uo, vo, ito = O.tvl1(I0, I1, **kw)
    if np.array_equal(it, ito):
        assert np.array_equal(u, uo) and np.array_equal(v, vo)
    else:
        assert np.abs(it - ito).max() <= 1
        epe = np.sqrt((u - uo) ** 2 + (v - vo) ** 2)
        assert epe.mean() < 1e-3
    r = O.tvl1_reference(I0, I1, **kw)
    if r is not None:
        epe = np.sqrt((u - r[0]) ** 2 + (v - r[1]) ** 2)
        assert epe.mean() < 1e-3


def test_tvl1_sequence_layout_and_end_copies(ctx):
    frames = synth.add_noise(synth.video(48, 40, 4, seed=9), 20, 10)[:, 0]
    kw = dict(nscales=3, warps=2)
    ff, bf = ctx.tvl1_sequence(frames, **kw)
    for i in range(3):
        u, v, _ = O.tvl1(frames[i], frames[i + 1], **kw)
        assert np.abs(ff[i, 0] - u).max() < 0.05 and np.abs(ff[i, 1] - v).max() < 0.05
        u, v, _ = O.tvl1(frames[i + 1], frames[i], **kw)
        assert np.abs(bf[i + 1, 0] - u).max() < 0.05 and np.abs(bf[i + 1, 1] - v).max() < 0.05
    assert np.array_equal(bf[0], bf[1]) and np.array_equal(ff[3], ff[2])


def test_dual_tvl1_dropin_symbol(ctx):
    """The reference's own entry point name and argument list."""
    import ctypes as C
    I0, I1 = _pair(48, 40, 3)
    u = np.zeros_like(I0)
    v = np.zeros_like(I0)
    A.lib().Dual_TVL1_optic_flow_multiscale(I0.ctypes.data, I1.ctypes.data, u.ctypes.data,
                                            v.ctypes.data, 48, 40, 0.25, 0.15, 0.3, 3, 0, 0.5, 3,
                                            0.01, False)
    uo, vo, _ = O.tvl1(I0, I1, nscales=3, warps=3)
    assert np.sqrt((u - uo) ** 2 + (v - vo) ** 2).mean() < 1e-3


def test_tvl1_matches_reference_vectors(ctx):
    """Against flows of the reference's compiled lib/tvl1flow committed under
    tests/golden (same criterion as against the oracle)."""
    from test_golden import tvl1_cases
    for i, I0, I1, ur, vr, kw in tvl1_cases():
        print("golden case", i, kw, flush=True)
        u, v, it = ctx.tvl1(I0, I1, **kw)
        _, _, ito = O.tvl1(I0, I1, **kw)
        if np.array_equal(it, ito):
            assert np.array_equal(u, ur) and np.array_equal(v, vr)
        else:
            assert np.sqrt((u - ur) ** 2 + (v - vr) ** 2).mean() < 1e-3
