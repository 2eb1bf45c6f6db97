"""The predictive block matching of a VBM3D_SEARCH build of the reference
(nldct.cpp:904-1130, 1368-1589), context option "vbm3d_search": neighbour
lists and both estimates bitwise equal to the CPU oracle (itself pinned
against the float64 statement in test_oracle_mini)."""
import numpy as np
import pytest

import oracle as O
from conftest import bm4dof_amd as A, synth
from test_gpu_denoise import _same, _smooth_flow

pytestmark = pytest.mark.gpu


@pytest.fixture
def pred(ctx):
    ctx.set_option("vbm3d_search", 1)
    O.set_predictive(True)
    yield ctx
    ctx.set_option("vbm3d_search", 0)
    O.set_predictive(False)


def _both(step, shape, sigma=20.0, wx=9, wxp=5, npf=2, dsub=None, **kw):
    F = shape[0]
    a = A.cli_params(step, sigma, shape, px=4, pt=4, wx=wx, np_=16, **kw)
    o = O.cli_params(step, sigma, F, px=4, pt=4, wx=wx, np_=16, **kw)
    for p in (a, o):
        p.sizeSearchWindowPred = wxp
        p.nSimilarPatchesPred = npf
        if dsub is not None:
            p.dsub = dsub
    return a, o


@pytest.mark.parametrize("channels,nparts,flow,quant", [(1, 2, True, False), (1, 4, False, True),
                                                       (3, 2, True, False), (3, 1, True, True)])
def test_predictive_neighbour_lists(pred, channels, nparts, flow, quant):
    ctx = pred
    clean = synth.video(64, 48, 8, channels=channels, seed=40 + channels)
    noisy = synth.add_noise(clean, 20, 41)
    if quant:   # integer samples: equal distances, libstdc++'s heap order decides
        noisy = np.clip(np.round(noisy), 0, 255).astype(np.float32)
    ff = bf = None
    if flow:
        ff, bf = _smooth_flow(8, 48, 64, 42), _smooth_flow(8, 48, 64, 43)
    for step, img in ((1, noisy), (2, clean)):
        a, o = _both(step, noisy.shape, npf=3, wxp=3)
        gi, gn, gr, gt = ctx.bm(a, img, ff, bf, nparts=nparts)
        oi, on, orf, ot = O.bm(o, img, ff, bf, nparts=nparts)
        assert np.array_equal(gr, orf) and np.array_equal(gt, ot)
        assert np.array_equal(gn, on), (step, int((gn != on).sum()))
        assert np.array_equal(gi, oi), (step, int((gi != oi).any(axis=1).sum()))


@pytest.mark.parametrize("channels", [1, 3])
def test_predictive_two_steps_parity(pred, channels):
    ctx = pred
    clean = synth.video(72, 56, 9, channels=channels, seed=50 + channels)
    noisy = synth.add_noise(clean, 20, 51)
    ff, bf = _smooth_flow(9, 56, 72, 52), _smooth_flow(9, 56, 72, 53)
    a1, o1 = _both(1, noisy.shape)
    a2, o2 = _both(2, noisy.shape)
    ob, of_, ogr = O.run(o1, o2, noisy, ff, bf, nparts=2)
    gb, gf, ggr = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    _same(gb, ob, "basic")
    _same(gf, of_, "final")
    assert np.allclose(ggr, ogr, rtol=0, atol=1e-4)
    # a different search than the exhaustive one
    ctx.set_option("vbm3d_search", 0)
    _, gf0, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    ctx.set_option("vbm3d_search", 1)
    assert not np.array_equal(gf0, gf)


def test_predictive_rejects_small_regions(pred):
    ctx = pred
    noisy = synth.add_noise(synth.video(40, 32, 6, seed=5), 20, 6)
    a1, _ = _both(1, noisy.shape, wxp=3, npf=10)
    a2, _ = _both(2, noisy.shape)
    with pytest.raises(A.BM4DOFError, match="nSimilarPatchesPred"):
        ctx.run(a1, a2, noisy, nparts=2)
