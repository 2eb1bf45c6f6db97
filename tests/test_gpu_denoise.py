"""GPU parity of the denoiser (libbm4dof.so through its C-ABI) against the
CPU oracle, on seeded synthetic inputs the oracle finishes in seconds.

Bars: block-matching neighbour lists identical (inputs have distinct
distances, so the GPU's total order and the reference's std::partial_sort
order agree); basic and final estimates within 1e-3 max-abs on the 0-255
scale and end-to-end PSNR within 0.01 dB.  The GPU path reproduces the
reference's float summation order in the distances, the forward transform,
the Wiener energy and the aggregation, so `_same` also asserts bitwise
equality: a reordering anywhere shows up as a failure here.
"""
import numpy as np
import pytest

import oracle as O
from conftest import bm4dof_amd as A, synth

pytestmark = pytest.mark.gpu

TOL = 1e-3


def _same(g, o, what=""):
    """Bitwise equal, NaN where the oracle has NaN (the reference's own
    arithmetic gives NaN where a step-2 group has an infinite weight, i.e.
    all its Wiener gains are 0, times a zero estimate)."""
    gn, on = np.isnan(g), np.isnan(o)
    assert np.array_equal(gn, on), (what, "NaN positions", int(gn.sum()), int(on.sum()))
    d = np.abs(g.astype(np.float64) - o)[~on]
    if d.size:
        assert d.max() <= TOL, (what, d.max())
    assert np.array_equal(g, o, equal_nan=True), (what, int((d > 0).sum()), d.max() if d.size else 0)


def _video(w, h, f, seed, sigma=20.0, channels=1):
    clean = synth.video(w, h, f, channels=channels, seed=seed)
    return clean, synth.add_noise(clean, sigma, seed + 1000)


def _smooth_flow(F, H, W, seed, amp=2.5):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    fl = np.zeros((F, 2, H, W), np.float32)
    for t in range(F):
        for c in range(2):
            a, b, ph = rng.uniform(-1, 1, 3)
            fl[t, c] = amp * np.sin(2 * np.pi * (a * xx / W + b * yy / H) + ph)
    return fl


def _params(shape, step, sigma=20.0, **kw):
    F = shape[0]
    return (A.cli_params(step, sigma, shape, **kw), O.cli_params(step, sigma, F, **kw))


VIDEO_CASES = [
    # (w, h, f, nparts, flow)
    (64, 48, 6, 2, False),
    (72, 40, 8, 1, True),
    (96, 64, 8, 4, True),
    (50, 38, 7, 6, True),
]


@pytest.mark.parametrize("w,h,f,nparts,flow", VIDEO_CASES)
def test_block_matching_neighbour_lists(ctx, w, h, f, nparts, flow):
    clean, noisy = _video(w, h, f, seed=w + f)
    ff = bf = None
    if flow:
        ff, bf = _smooth_flow(f, h, w, 1), _smooth_flow(f, h, w, 2)
    a1, o1 = _params(noisy.shape, 1, px=4, pt=4, wx=11, np_=32)
    gi, gn, gr, gt = ctx.bm(a1, noisy, ff, bf, nparts=nparts)
    oi, on, orf, ot = O.bm(o1, noisy, ff, bf, nparts=nparts)
    assert np.array_equal(gr, orf) and np.array_equal(gt, ot)
    assert np.array_equal(gn, on)
    assert np.array_equal(gi, oi)
    # step 2 distances are taken on the basic image (all channels)
    basic = (clean + np.random.default_rng(0).normal(0, 3, clean.shape)).astype(np.float32)
    a2, o2 = _params(noisy.shape, 2, px=4, pt=4, wx=11, np_=32)
    gi, gn, _, _ = ctx.bm(a2, basic, ff, bf, nparts=nparts)
    oi, on, _, _ = O.bm(o2, basic, ff, bf, nparts=nparts)
    assert np.array_equal(gn, on) and np.array_equal(gi, oi)


@pytest.mark.parametrize("mode,channels,flow", [(2, 1, True), (1, 1, False), (0, 1, True), (2, 3, True),
                                               (0, 3, False)])
def test_quantized_video_ties_follow_the_reference(ctx, mode, channels, flow):
    """Integer-valued noisy video (as 8-bit frames read from PNG files): every
    distance is an integer and equal distances are common.  The reference
    leaves equal distances in std::partial_sort's heap order; each matcher
    lists such groups and matches them again in that order, so the
    neighbour lists and both estimates equal the oracle's libstdc++ order
    (tie_mode 0) bit for bit."""
    clean, noisy = _video(56, 40, 7, seed=11 + mode + channels, channels=channels)
    noisy = np.clip(np.round(noisy), 0, 255).astype(np.float32)
    ff = bf = None
    if flow:
        ff, bf = _smooth_flow(7, 40, 56, 13), _smooth_flow(7, 40, 56, 14)
    a1, o1 = _params(noisy.shape, 1, px=4, pt=4, wx=9, np_=16)
    a2, o2 = _params(noisy.shape, 2, px=4, pt=4, wx=9, np_=16)
    try:
        ctx.set_option("fast_bm", mode)
        gi, gn, _, _ = ctx.bm(a1, noisy, ff, bf, nparts=2)
        oi, on, _, _ = O.bm(o1, noisy, ff, bf, nparts=2)
        assert np.array_equal(gn, on) and np.array_equal(gi, oi)
        # the total order alone would differ here: ties do occur
        ti, tn, _, _ = O.bm(o1, noisy, ff, bf, nparts=2, tie_mode=1)
        assert not (np.array_equal(ti, oi) and np.array_equal(tn, on))
        ob, of_, _ = O.run(o1, o2, noisy, ff, bf, nparts=2)
        gb, gf, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    finally:
        ctx.set_option("fast_bm", 2)
    _same(gb, ob, "basic")
    _same(gf, of_, "final")


@pytest.mark.parametrize("w,h,f,nparts,flow", VIDEO_CASES)
def test_video_two_step_parity(ctx, w, h, f, nparts, flow):
    clean, noisy = _video(w, h, f, seed=3 * w + f)
    ff = bf = None
    if flow:
        ff, bf = _smooth_flow(f, h, w, 3), _smooth_flow(f, h, w, 4)
    a1, o1 = _params(noisy.shape, 1, px=4, pt=4, wx=11, np_=32)
    a2, o2 = _params(noisy.shape, 2, px=4, pt=4, wx=11, np_=32)
    ob, of_, ogr = O.run(o1, o2, noisy, ff, bf, nparts=nparts)
    gb, gf, ggr = ctx.run(a1, a2, noisy, ff, bf, nparts=nparts)
    _same(gb, ob, 'basic')
    assert ggr == pytest.approx(ogr, abs=0)
    # step 2 alone from the CPU basic estimate (DENO_ONLY, -px1 0 -b)
    a0 = A.cli_params(1, 20, noisy.shape, px=0)
    _, gf2, _ = ctx.run(a0, a2, noisy, ff, bf, basic=ob, nparts=nparts)
    _same(gf2, of_, 'final from CPU basic')
    # end to end
    _same(gf, of_, 'final')
    dpsnr = O.psnr(gf, clean)[0] - O.psnr(of_, clean)[0]
    assert abs(dpsnr) <= 0.01


@pytest.mark.parametrize("px,pt,np_,frames", [(4, 4, 32, 8), (8, 1, 16, 1), (4, 4, 8, 6), (8, 1, 32, 3)])
def test_register_and_lds_filters_agree(ctx, px, pt, np_, frames):
    """The register-resident filter (one channel, 64-sample patches) and the
    general LDS filter give bitwise equal basic and final estimates; so does
    the register filter without the per-anchor coefficient tables
    (coef_budget 0: every group transforms its patches itself)."""
    clean, noisy = _video(80, 56, frames, seed=10 * px + np_)
    a1 = A.cli_params(1, 20, noisy.shape, px=px, pt=pt, wx=9, np_=np_)
    a2 = A.cli_params(2, 20, noisy.shape, px=px, pt=pt, wx=9, np_=np_)
    try:
        ctx.set_option("fast_filter", 0)
        b0, f0, _ = ctx.run(a1, a2, noisy, nparts=2)
    finally:
        ctx.set_option("fast_filter", 1)
    b1, f1, _ = ctx.run(a1, a2, noisy, nparts=2)
    assert np.array_equal(b0, b1) and np.array_equal(f0, f1)
    try:
        ctx.set_option("coef_budget", 0)
        b2, f2, _ = ctx.run(a1, a2, noisy, nparts=2)
    finally:
        ctx.set_option("coef_budget", 96 << 30)
    assert np.array_equal(b2, b1) and np.array_equal(f2, f1)
    with pytest.raises(A.BM4DOFError, match="unknown option"):
        ctx.set_option("no_such_option", 1)


@pytest.mark.parametrize("w,h,f,nparts,px,pt", [(96, 64, 8, 4, 4, 4), (64, 48, 6, 2, 4, 4),
                                                 (80, 72, 1, 4, 8, 1), (50, 38, 7, 6, 4, 4),
                                                 (72, 56, 9, 2, 4, 2)])
def test_chunked_and_banded_aggregation_agree(ctx, w, h, f, nparts, px, pt):
    """The chunked aggregation (per-pixel masks of sorted row chunks) and the
    banded one fold the same rows in the same order: bitwise equal output,
    also where a pixel block straddles four tiles (nparts=4: 2 x 2 tiles)."""
    clean, noisy = _video(w, h, f, seed=w * h + f)
    ff = bf = None
    if f > 1:
        ff, bf = _smooth_flow(f, h, w, 5), _smooth_flow(f, h, w, 6)
    a1 = A.cli_params(1, 20, noisy.shape, px=px, pt=pt, wx=9, np_=16)
    a2 = A.cli_params(2, 20, noisy.shape, px=px, pt=pt, wx=9, np_=16)
    try:
        ctx.set_option("agg_chunked", 0)
        b0, f0, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=nparts)
    finally:
        ctx.set_option("agg_chunked", 1)
    b1, f1, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=nparts)
    assert np.array_equal(b0, b1) and np.array_equal(f0, f1)


@pytest.mark.parametrize("w,h,f,c,nparts,px,pt", [(96, 64, 8, 1, 4, 4, 4), (50, 38, 7, 1, 6, 4, 4),
                                                   (80, 72, 1, 1, 4, 8, 1), (72, 56, 9, 1, 2, 4, 2),
                                                   (61, 45, 6, 3, 2, 4, 4), (64, 48, 5, 4, 4, 4, 4),
                                                   (88, 40, 12, 1, 2, 4, 3)])
def test_binned_and_chunked_aggregation_agree(ctx, w, h, f, c, nparts, px, pt):
    """The binned aggregation (row lists per 8 x 8 x 4 bin from a stable
    radix sort of (bin, row) pairs, four frames folded per wave) and the
    per-frame chunked one fold the same rows in the same order: bitwise equal
    output for gray and colour videos, frame counts that do not fill the last
    bin, blocks straddling 2 x 2 tiles."""
    clean, noisy = _video(w, h, f, seed=w * h + f + c, channels=c)
    ff = bf = None
    if f > 1:
        ff, bf = _smooth_flow(f, h, w, 15), _smooth_flow(f, h, w, 16)
    a1 = A.cli_params(1, 20, noisy.shape, px=px, pt=pt, wx=9, np_=16)
    a2 = A.cli_params(2, 20, noisy.shape, px=px, pt=pt, wx=9, np_=16)
    try:
        ctx.set_option("agg_binned", 0)
        b0, f0, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=nparts)
    finally:
        ctx.set_option("agg_binned", 1)
    b1, f1, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=nparts)
    # bins take the 64-sample patches (4x4x4, 8x8x1); other shapes keep the
    # per-frame blocks
    assert ctx.stats()["agg_bt"] == (4 if (px, pt) in ((4, 4), (8, 1)) else 1)
    assert np.array_equal(b0, b1, equal_nan=True) and np.array_equal(f0, f1, equal_nan=True)


@pytest.mark.parametrize("blocks", [1, 3, 64])
def test_bin_lists_with_few_second_stream_blocks(ctx, blocks):
    """The bin-list kernels loop over their work when the second stream's
    grid is capped (option side_blocks): one block, three, 64 blocks give
    bitwise the output of the default grid, colour and gray."""
    for c in (1, 3):
        clean, noisy = _video(88, 64, 9, seed=500 + blocks + c, channels=c)
        ff, bf = _smooth_flow(9, 64, 88, 21), _smooth_flow(9, 64, 88, 22)
        a1 = A.cli_params(1, 20, noisy.shape, px=4, pt=4, wx=9, np_=16)
        a2 = A.cli_params(2, 20, noisy.shape, px=4, pt=4, wx=9, np_=16)
        b0, f0, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
        try:
            ctx.set_option("side_blocks", blocks)
            b1, f1, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
        finally:
            ctx.set_option("side_blocks", 2048)
        assert ctx.stats()["agg_bt"] == 4
        assert np.array_equal(b0, b1, equal_nan=True) and np.array_equal(f0, f1, equal_nan=True)


@pytest.mark.parametrize("px,pt,flow,nparts,wx,flat", [(4, 4, True, 4, 11, False), (8, 1, False, 2, 11, False),
                                                       (4, 4, False, 6, 11, False), (4, 4, True, 2, 11, True),
                                                       (8, 1, False, 1, 21, True), (8, 1, False, 2, 39, False),
                                                       (4, 4, False, 2, 5, False)])
def test_block_matching_kernels_agree(ctx, px, pt, flow, nparts, wx, flat):
    """Neighbour lists of the register-window and the LDS-staged one-wave
    block matching equal those of the general kernel, in both steps.  `flat`
    puts a constant region in the image: many candidates at distance 0, which
    sends the register-window selection to its radix fallback."""
    f = 8 if pt > 1 else 1
    clean, noisy = _video(72, 56, f, seed=px + 7 * nparts + wx)
    if flat:
        noisy[:, :, :, :30] = 77.0
        clean[:, :, :, :30] = 77.0
    ff = bf = None
    if flow:
        ff, bf = _smooth_flow(f, 56, 72, 7), _smooth_flow(f, 56, 72, 8)
    for step, img in ((1, noisy), (2, clean)):
        a = A.cli_params(step, 20, noisy.shape, px=px, pt=pt, wx=wx, np_=32)
        res = {}
        try:
            for mode in (0, 1, 2):
                ctx.set_option("fast_bm", mode)
                res[mode] = ctx.bm(a, img, ff, bf, nparts=nparts)
        finally:
            ctx.set_option("fast_bm", 2)
        for mode in (1, 2):
            for u, v in zip(res[0], res[mode]):
                assert np.array_equal(u, v), (step, mode)


def test_image_config0_bior_and_kaiser(ctx):
    """Single grayscale image, px=8 (bior1.5 + Kaiser window in step 1, DCT in
    step 2), np=16, 21x21 search, beta1 = 3."""
    clean = synth.image(96, 80, seed=100)
    noisy = synth.add_noise(clean, 20, 200)
    a1, o1 = _params(noisy.shape, 1, px=8, pt=1, wx=21, np_=16, beta=3.0)
    a2, o2 = _params(noisy.shape, 2, px=8, pt=1, wx=21, np_=16)
    ob, of_, _ = O.run(o1, o2, noisy, nparts=2)
    gb, gf, _ = ctx.run(a1, a2, noisy, nparts=2)
    _same(gb, ob, 'basic')
    _, gf2, _ = ctx.run(A.cli_params(1, 20, noisy.shape, px=0), a2, noisy, basic=ob, nparts=2)
    _same(gf2, of_, 'final from CPU basic')
    assert O.psnr(gf, clean)[0] > O.psnr(noisy, clean)[0] + 5


def test_color_image_parity(ctx):
    clean = synth.video(48, 40, 1, channels=3, seed=21)
    noisy = synth.add_noise(clean, 20, 22)
    a1, o1 = _params(noisy.shape, 1, px=8, pt=1, wx=11, np_=16)
    a2, o2 = _params(noisy.shape, 2, px=8, pt=1, wx=11, np_=16)
    ob, of_, _ = O.run(o1, o2, noisy, nparts=2)
    gb, gf, _ = ctx.run(a1, a2, noisy, nparts=2)
    _same(gb, ob, 'basic')
    _same(gf, of_, 'final')
    # The returned basic estimate went RGB->YUV->RGB once more after step 2
    # (nldct.cpp:423, 467), so step 2 alone from it differs from the
    # two-step run by rounding; compare it with the oracle's step 2 alone.
    o0 = O.cli_params(1, 20, 1, px=0)
    _, of2, _ = O.run(o0, o2, noisy, basic=ob, nparts=2)
    _, gf2, _ = ctx.run(A.cli_params(1, 20, noisy.shape, px=0), a2, noisy, basic=ob, nparts=2)
    _same(gf2, of2, 'final from CPU basic')


@pytest.mark.parametrize("channels,flow", [(3, False), (3, True), (4, True)])
def test_color_video_parity(ctx, channels, flow):
    """Colour video: opponent transform, step-1 matching on Y, step-2 on all
    channels, and the mask clear with the multichannel index
    (nldct.cpp:2325-2329) that skips grid positions of frames C*t."""
    clean, noisy = _video(40, 32, 7, seed=30 + channels, channels=channels)
    a1, o1 = _params(noisy.shape, 1, px=4, pt=2, wx=7, np_=8)
    a2, o2 = _params(noisy.shape, 2, px=4, pt=2, wx=7, np_=8)
    ff = bf = None
    if flow:
        ff = _smooth_flow(7, 32, 40, 5)
        bf = _smooth_flow(7, 32, 40, 6)
    ob, of_, ogr = O.run(o1, o2, noisy, ff, bf, nparts=2)
    gb, gf, ggr = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    _same(gb, ob, "basic")
    _same(gf, of_, "final")
    assert np.allclose(ggr, ogr, rtol=0, atol=1e-4)
    # the skipped positions: fewer groups than the same video in gray
    g1 = noisy[:, :1].copy()
    _, _, gr_gray = ctx.run(A.cli_params(1, 20, g1.shape, px=4, pt=2, wx=7, np_=8),
                            A.cli_params(2, 20, g1.shape, px=4, pt=2, wx=7, np_=8), g1, ff, bf, nparts=2)
    assert ggr[0] < gr_gray[0] and ggr[1] < gr_gray[1]


@pytest.mark.parametrize("channels,px,pt,flow,flat", [(3, 4, 4, True, False), (3, 4, 4, False, True),
                                                      (4, 8, 1, False, True)])
def test_color_register_paths_parity(ctx, channels, px, pt, flow, flat):
    """Colour video through the register-resident filters (one launch per
    channel, the group weight counting every channel) and the register-window
    matching (step-2 distances summed over the channels): bitwise equal to
    the CPU oracle, to the general kernels and to the register filters
    without coefficient tables (coef_budget 0).  `flat` (RGB, no transform)
    puts a region constant in every channel, which takes the equality
    early-out, beside one constant in the first channel only, which must
    not."""
    f = 7 if pt > 1 else 2
    clean, noisy = _video(48, 40, f, seed=60 + channels + px, channels=channels)
    if flat:
        for c in range(channels):
            noisy[:, c, :, :14] = 40.0 + 30 * c
        noisy[:, 0, 24:, 20:] = 90.0
    ff = bf = None
    if flow:
        ff, bf = _smooth_flow(f, 40, 48, 3), _smooth_flow(f, 40, 48, 4)
    a1, o1 = _params(noisy.shape, 1, px=px, pt=pt, wx=7, np_=16)
    a2, o2 = _params(noisy.shape, 2, px=px, pt=pt, wx=7, np_=16)
    if flat:
        for p in (a1, o1, a2, o2):
            p.colorSpace = 0
    ob, of_, ogr = O.run(o1, o2, noisy, ff, bf, nparts=2)
    gb, gf, ggr = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    _same(gb, ob, "basic")
    _same(gf, of_, "final")
    try:
        ctx.set_option("fast_filter", 0)
        ctx.set_option("fast_bm", 0)
        b0, f0, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    finally:
        ctx.set_option("fast_filter", 1)
        ctx.set_option("fast_bm", 2)
    assert np.array_equal(b0, gb, equal_nan=True) and np.array_equal(f0, gf, equal_nan=True)
    try:
        ctx.set_option("coef_budget", 0)
        b2, f2, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    finally:
        ctx.set_option("coef_budget", 96 << 30)
    assert np.array_equal(b2, gb, equal_nan=True) and np.array_equal(f2, gf, equal_nan=True)


@pytest.mark.parametrize("cs1,cs2", [(1, 0), (0, 1)])
def test_color_steps_in_different_colour_spaces(ctx, cs1, cs2):
    """Steps in different colour spaces (one YUV, one RGB) filter different
    noisy planes, so step 2 computes its own noisy coefficient table instead
    of reading step 1's: bitwise equal to the CPU oracle, which transforms
    around each step by its own parameter (nldct.cpp:363, 401, 422-423)."""
    clean, noisy = _video(48, 40, 7, seed=300 + cs1, channels=3)
    ff, bf = _smooth_flow(7, 40, 48, 8), _smooth_flow(7, 40, 48, 9)
    a1, o1 = _params(noisy.shape, 1, px=4, pt=4, wx=7, np_=16)
    a2, o2 = _params(noisy.shape, 2, px=4, pt=4, wx=7, np_=16)
    a1.colorSpace = o1.colorSpace = cs1
    a2.colorSpace = o2.colorSpace = cs2
    ob, of_, _ = O.run(o1, o2, noisy, ff, bf, nparts=2)
    gb, gf, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    _same(gb, ob, "basic")
    _same(gf, of_, "final")


@pytest.mark.parametrize("channels,flow", [(3, True), (4, False)])
def test_color_block_matching_kernels_agree(ctx, channels, flow):
    """Step-2 matching on colour video sums the distances of all channels
    in channel order: the register-window kernel's neighbour lists equal
    the general kernel's."""
    f = 8
    clean, _ = _video(64, 48, f, seed=90 + channels, channels=channels)
    clean[:, :, :20, :20] = 60.0
    ff = bf = None
    if flow:
        ff, bf = _smooth_flow(f, 48, 64, 9), _smooth_flow(f, 48, 64, 10)
    a = A.cli_params(2, 20, clean.shape, px=4, pt=4, wx=11, np_=32)
    res = {}
    try:
        for mode in (0, 2):
            ctx.set_option("fast_bm", mode)
            res[mode] = ctx.bm(a, clean, ff, bf, nparts=2)
    finally:
        ctx.set_option("fast_bm", 2)
    for u, v in zip(res[0], res[2]):
        assert np.array_equal(u, v)


def test_order_invariance_and_np1(ctx):
    clean, noisy = _video(40, 32, 6, seed=77)
    for kw in (dict(order_inv=True), dict(np_=1)):
        a1, o1 = _params(noisy.shape, 1, px=4, pt=4, wx=7, **dict(dict(np_=16), **kw))
        a2, o2 = _params(noisy.shape, 2, px=4, pt=4, wx=7, **dict(dict(np_=16), **kw))
        ob, of_, _ = O.run(o1, o2, noisy, nparts=2)
        gb, gf, _ = ctx.run(a1, a2, noisy, nparts=2)
        _same(gb, ob, kw)
        _, gf2, _ = ctx.run(A.cli_params(1, 20, noisy.shape, px=0), a2, noisy, basic=ob, nparts=2)
        _same(gf2, of_, kw)


def test_short_video_and_constant(ctx):
    # fewer frames than the temporal patch: no reference patches, all-zero
    # output (nldct.cpp:537, 2446)
    _, noisy = _video(32, 24, 3, seed=5)
    a1, o1 = _params(noisy.shape, 1, px=4, pt=4, wx=7, np_=8)
    a2, o2 = _params(noisy.shape, 2, px=4, pt=4, wx=7, np_=8)
    gb, gf, ggr = ctx.run(a1, a2, noisy, nparts=2)
    ob, of_, ogr = O.run(o1, o2, noisy, nparts=2)
    assert np.array_equal(gb, ob) and np.array_equal(gf, of_) and not gb.any()
    assert ggr == (0.0, 0.0)
    # constant video: every distance is 0, so the reference's heap order
    # keeps the first patches of each window, not the reference patch; its
    # mask clears then skip grid positions, some pixels stay uncovered (and
    # are set to 0, nldct.cpp:2446) and every group takes the equality
    # early-out with stale weights.  Equal to the oracle, groups ratio
    # included (under the total order the output would equal the input).
    v = np.full((6, 1, 24, 32), 42.0, np.float32)
    gb, gf, ggr = ctx.run(a1, a2, v, nparts=2)
    ob, of_, ogr = O.run(o1, o2, v, nparts=2)
    _same(gb, ob, "basic")
    _same(gf, of_, "final")
    assert ggr == pytest.approx(ogr, abs=0)
    assert (gb == 0).any()


def test_partly_flat_video_stale_weights(ctx):
    """A video whose left part is noise-free: groups of identical patches take
    the early-out and reuse the previous group's weights (nldct.cpp:1998)."""
    clean, noisy = _video(48, 32, 6, seed=31)
    noisy[:, :, :, :20] = 60.0
    a1, o1 = _params(noisy.shape, 1, px=4, pt=4, wx=7, np_=8)
    a2, o2 = _params(noisy.shape, 2, px=4, pt=4, wx=7, np_=8)
    ob, of_, _ = O.run(o1, o2, noisy, nparts=2)
    gb, gf, _ = ctx.run(a1, a2, noisy, nparts=2)
    _same(gb, ob, 'basic')
    _, gf2, _ = ctx.run(A.cli_params(1, 20, noisy.shape, px=0), a2, noisy, basic=ob, nparts=2)
    _same(gf2, of_, 'final from CPU basic')


def test_errors_match_reference_messages(ctx):
    _, noisy = _video(32, 24, 6, seed=5)
    a1 = A.cli_params(1, 20, noisy.shape, px=0)
    a2 = A.cli_params(2, 20, noisy.shape, px=0)
    with pytest.raises(A.BM4DOFError, match="Both patch sizes are zero"):
        ctx.run(a1, a2, noisy)
    two = np.zeros((6, 2, 24, 32), np.float32)
    b1 = A.cli_params(1, 20, two.shape, px=4, pt=4)
    with pytest.raises(A.BM4DOFError, match="Wrong number of channels"):
        ctx.run(b1, a2, two)
    bior = A.cli_params(1, 20, (1, 1, 24, 32), px=8)
    bior.sizePatch = 7
    with pytest.raises(A.BM4DOFError, match="bior1_5"):
        ctx.run(bior, a2, noisy[:1])


def test_deterministic_and_full_size_properties(ctx):
    """960x540x20, sigma 20 (the benchmark workload): two runs are bitwise
    equal, the output is finite and the PSNR gain is large."""
    clean = synth.video(960, 540, 20, seed=101)
    noisy = synth.add_noise(clean, 20, 201)
    a1 = A.cli_params(1, 20, noisy.shape, px=4, pt=4, wx=11, np_=32)
    a2 = A.cli_params(2, 20, noisy.shape, px=4, pt=4, wx=11, np_=32)
    b1, f1, gr = ctx.run(a1, a2, noisy, nparts=2)
    b2, f2, _ = ctx.run(a1, a2, noisy, nparts=2)
    assert np.array_equal(b1, b2) and np.array_equal(f1, f2)
    assert np.isfinite(f1).all()
    assert O.psnr(f1, clean)[0] > O.psnr(noisy, clean)[0] + 15
    assert O.psnr(f1, clean)[0] > O.psnr(b1, clean)[0]
    assert gr[0] > 9 and gr[1] > 21
