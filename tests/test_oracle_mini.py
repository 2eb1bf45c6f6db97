"""Pins of the CPU oracle against a second, independent statement of the
reference (tests/nldct_mini.py: float64 numpy written from the reference
text, libstdc++'s own std::partial_sort for the candidate order).

The parts pinned here are the ones no closed form fixes: the flow trajectory
and its window shift (src/core/nldct.cpp:1219-1259), the step-2 distance over
all channels of the basic image (:1723-1737), the reference grid of interior
tiles (:513-553) and the write-back of tile interiors
(src/utils/lib_videot.hpp:1102-1163), the mask clear at the first neighbour
(:2325-2329, with the colour index of :2328), the stale aggregation weights of
equality early-outs (:749, :816, :1986-1999, :2222-2235), the orderInvariance
filters (:1865-1880, :2101-2115), the group ratios (:403-404, :471-472) and
the bior1.5 step 1 at 8x8 patches.

Inputs are integer-valued (so equal distances really tie) and sigma is 20.3
(so that no transform coefficient of an integer patch sits on the hard
threshold).
Each comparison is checked to be sensitive: the same run with one planted
misreading in the float64 statement must disagree with the oracle.
"""
import numpy as np
import pytest

import oracle as O
import nldct_mini as M
from conftest import synth

SIG = 20.3


def _int_video(w, h, f, c=1, seed=3, sigma=20.0, hi=255):
    clean = synth.video(w, h, f, channels=c, seed=seed, n_shapes=6)
    noisy = synth.add_noise(clean, sigma, seed + 1000)
    return np.clip(np.round(noisy * (hi / 255.0)), 0, hi).astype(np.float32)


def _flows(f, h, w, seed, scale=1.0, halves=True, push=0.0):
    """Forward/backward flows of (F, 2, H, W): random per pixel, with exact
    halves (to tell the roundings apart), quarters, and optionally a push
    towards the frame edges (clamping)."""
    rng = np.random.default_rng(seed)
    ff = np.zeros((f, 2, h, w), np.float32)
    bf = np.zeros((f, 2, h, w), np.float32)
    steps = np.array([-2.5, -1.5, -1, -0.5, -0.25, 0, 0.25, 0.5, 1, 1.5, 2.5] if halves
                     else [-2, -1, 0, 1, 2], np.float32)
    for fl, sgn in ((ff, 1.0), (bf, -1.0)):
        fl[:] = rng.choice(steps, size=fl.shape) * scale
        if push:
            yy, xx = np.mgrid[0:h, 0:w]
            fl[:, 0] += sgn * push * np.sign(xx - w / 2)
            fl[:, 1] += sgn * push * np.sign(yy - h / 2)
    return ff, bf


def _params(step, F, **kw):
    kw.setdefault("px", 4)
    kw.setdefault("pt", 4 if F > 1 else 1)
    kw.setdefault("wx", 5)
    kw.setdefault("np_", 16)
    return O.cli_params(step, SIG, F, **kw)


def _bm_lists_equal(p, image, ff, bf, nparts, mistake=None):
    """Neighbour lists of every grid point of every tile: oracle vs the
    float64 statement.  Returns (number of groups, number that differ)."""
    idx, nsim, ref, tile = O.bm(p, image, ff, bf, nparts=nparts, tie_mode=0)
    mini = M.Mini(p, p, mistake=mistake)
    step = 1 if p.isFirstStep else 2
    groups = mini.bm_all(step, image, ff, bf, nparts=nparts)
    if len(groups) != len(ref):
        return len(ref), max(len(ref), len(groups))
    bad = 0
    for g, (ti, r, lst) in enumerate(groups):
        if (ti != tile[g] or r != ref[g] or nsim[g] != len(lst)
                or list(idx[g, :nsim[g]]) != lst):
            bad += 1
    return len(groups), bad


# ------------------------------------------------------ flow trajectory

@pytest.mark.parametrize("nparts", [1, 4])
def test_trajectory_constant_integer_flow(nparts):
    """Constant integer motion (+2, -1) forward and (-2, +1) backward: the
    window moves along it frame by frame (nldct.cpp:1219-1236)."""
    v = _int_video(30, 26, 9, seed=5)
    F, _, H, W = v.shape
    ff = np.zeros((F, 2, H, W), np.float32)
    ff[:, 0], ff[:, 1] = 2, -1
    bf = -ff
    p = _params(1, F)
    n, bad = _bm_lists_equal(p, v, ff, bf, nparts)
    assert n > 50 and bad == 0


@pytest.mark.parametrize("nparts", [1, 2, 4])
def test_trajectory_subpixel_flow_with_halves_and_clamping(nparts):
    """Per-pixel flows with exact halves (half away from zero), quarters and
    a push past the frame edges (clamp to the tile, then the window shift,
    nldct.cpp:1227-1259); read at the previous trajectory point of the
    previous frame, forward and backward."""
    v = _int_video(30, 26, 9, seed=6)
    F, _, H, W = v.shape
    ff, bf = _flows(F, H, W, seed=11, push=4.0)
    p = _params(1, F)
    n, bad = _bm_lists_equal(p, v, ff, bf, nparts)
    assert n > 50 and bad == 0


@pytest.mark.parametrize("mistake", ["truncate", "round_half_even", "flow_at_reference",
                                     "flow_of_target_frame", "clamp_to_last_patch"])
def test_trajectory_pins_are_sensitive(mistake):
    v = _int_video(30, 26, 9, seed=6)
    F, _, H, W = v.shape
    ff, bf = _flows(F, H, W, seed=11, push=4.0)
    p = _params(1, F)
    n, bad = _bm_lists_equal(p, v, ff, bf, 4, mistake=mistake)
    assert bad > 0, mistake


# ----------------------------------------- step-2 distance over channels

@pytest.mark.parametrize("nparts", [1, 4])
def test_step2_distance_sums_all_channels(nparts):
    """Step 2 matches on the basic image summed over every channel
    (nldct.cpp:1723-1737); step 1 on channel 0 only (:1262-1274)."""
    v = _int_video(30, 26, 9, c=3, seed=7, hi=127)
    F, _, H, W = v.shape
    ff, bf = _flows(F, H, W, seed=12)
    for step in (1, 2):
        p = _params(step, F)
        n, bad = _bm_lists_equal(p, v, ff, bf, nparts)
        assert n > 50 and bad == 0, step
    p = _params(2, F)
    assert _bm_lists_equal(p, v, ff, bf, nparts, mistake="step2_channel0")[1] > 0


# ------------------------------------------------- tiles and their grid

@pytest.mark.parametrize("nparts", [2, 4, 6, 8])
def test_interior_tile_grid(nparts):
    """Grid of tiles with interior borders on every side (2x2, 3x2, 4x2
    tilings): origins and ends shifted by px-1+sWx/2 at interior borders,
    the last row/column only at true image edges (nldct.cpp:513-553)."""
    v = _int_video(44, 40, 6, seed=8)
    F = v.shape[0]
    p = _params(1, F, wx=3, np_=1)
    n, bad = _bm_lists_equal(p, v, None, None, nparts)
    assert n > 50 and bad == 0
    assert _bm_lists_equal(p, v, None, None, nparts, mistake="border_origin")[1] > 0


# ------------------------------------------------------- whole pipeline

def _maxdiff(a, b):
    """Largest difference, NaN counted as a difference unless both are NaN."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both = np.isnan(a) & np.isnan(b)
    d = np.abs(a - b)
    d[both] = 0
    d[np.isnan(d)] = np.inf
    return float(d.max())


def _compare_run(p1, p2, v, ff=None, bf=None, nparts=4, mistake=None, bior=None, tol=1e-3,
                 basic=None):
    """Oracle runNlBayes vs the float64 statement, step by step.  Step 2 of
    the float64 statement starts from the oracle's basic estimate (or from
    `basic` when step 1 is skipped), so that both match on the same image.
    Returns (max |basic diff|, max |final diff|, (step-1 and step-2
    statements))."""
    ob, of_, orat = O.run(p1, p2, v, ff, bf, basic=basic, nparts=nparts)
    F = v.shape[0]
    none = O.cli_params(2, SIG, F, px=0)
    m1 = M.Mini(p1, none, bior_bases=bior, mistake=mistake)
    db = 0.0
    if p1.sizePatch:
        mb, _, r1 = m1.run(v, ff, bf, nparts=nparts)
        db = _maxdiff(mb, ob)
        if mistake is None:
            assert abs(r1[0] - orat[0]) <= 1e-6 * max(1.0, orat[0])
    m2 = M.Mini(none, p2, bior_bases=bior, mistake=mistake)
    df = 0.0
    if p2.sizePatch:
        _, mf, r2 = m2.run(v, ff, bf, basic=ob, nparts=nparts)
        df = _maxdiff(mf, of_)
        if mistake is None:
            assert abs(r2[1] - orat[1]) <= 1e-6 * max(1.0, orat[1])
    if mistake is None:
        assert db < tol and df < tol, (db, df)
    return db, df, (m1, m2)


@pytest.mark.parametrize("nparts", [1, 2, 4])
def test_two_steps_with_flow_match_float64_statement(nparts):
    v = _int_video(30, 26, 9, seed=9)
    F, _, H, W = v.shape
    ff, bf = _flows(F, H, W, seed=13, push=2.0)
    _compare_run(_params(1, F), _params(2, F), v, ff, bf, nparts=nparts)


def test_colour_two_steps_in_opponent_space():
    """3-channel video in the opponent colour space: step 1 matches on Y,
    filters every channel and aggregates with one weight; step 2 matches on
    all channels of the basic image; the transforms around both steps."""
    v = _int_video(28, 24, 8, c=3, seed=10, hi=127)
    F, _, H, W = v.shape
    ff, bf = _flows(F, H, W, seed=14)
    p1, p2 = _params(1, F), _params(2, F)
    assert p1.colorSpace == 1
    _compare_run(p1, p2, v, ff, bf, nparts=2)


def _flat_video(c=1, seed=15, hi=255):
    """Integer video whose top-left part is one constant value: runs of
    identical patches (equal distances, equality early-outs)."""
    v = _int_video(32, 28, 8, c=c, seed=seed, hi=hi)
    v[:, :, :14, :18] = 40.0
    return v


def _flat_case(step, c, mistake=None):
    """Step 1 on a flat-region video, or step 2 alone (the basic estimate an
    input) on a noisy and a basic video that share the flat region."""
    hi = 255 if c == 1 else 100
    v = _flat_video(c=c, hi=hi)
    F = v.shape[0]
    if step == 1:
        p1, p2 = _params(1, F), O.cli_params(2, SIG, F, px=0)
        basic = None
    else:
        p1, p2 = O.cli_params(1, SIG, F, px=0), _params(2, F)
        basic = _flat_video(c=c, seed=25, hi=hi)
    p1.colorSpace = p2.colorSpace = 0  # RGB: channels as given (exact integers)
    return _compare_run(p1, p2, v, nparts=2, basic=basic, mistake=mistake)


@pytest.mark.parametrize("step", [1, 2])
@pytest.mark.parametrize("c", [1, 3])
def test_mask_clear_at_first_neighbour_and_stale_weights(step, c):
    """Where patches are identical, libstdc++'s partial_sort can put another
    patch first; its 4-argument index clears the processing mask there, so a
    later grid point is skipped (nldct.cpp:2325-2329; for 3 channels the index
    lands in frame 3t of the 1-channel mask).  Groups of identical patches take
    the equality early-out and aggregate with the weights the previous groups
    left, entry by entry (nldct.cpp:1986-1999, 2222-2235)."""
    _, _, ms = _flat_case(step, c)
    m = ms[step - 1]
    assert m.skipped[step] > 0 and m.stale[step] > 0, (m.skipped, m.stale)


@pytest.mark.parametrize("step", [1, 2])
@pytest.mark.parametrize("mistake", ["clear_reference", "fresh_weight", "clear_one_channel_index"])
def test_mask_and_weight_pins_are_sensitive(step, mistake):
    c = 3 if mistake == "clear_one_channel_index" else 1
    db, df, _ = _flat_case(step, c, mistake=mistake)
    assert max(db, df) > 1e-3, (mistake, db, df)


def test_order_invariance_filters():
    """-order-inv: per-coefficient Wiener gains pooled over the group's
    non-DC Haar rows, the DC with betaMean (nldct.cpp:1865-1880,
    2101-2115)."""
    v = _int_video(30, 26, 9, seed=16)
    F = v.shape[0]
    p1, p2 = _params(1, F, order_inv=True), _params(2, F, order_inv=True)
    p1.betaMean = p2.betaMean = 1.5
    # groups of one patch make the pooled variance 0/0 (NaN weights) in the
    # reference; keep every group at 16 patches
    p1.tau = p2.tau = 1e4
    db, df, _ = _compare_run(p1, p2, v, nparts=2)
    db2, _, _ = _compare_run(p1, p2, v, nparts=2, mistake="oi_no_dc_bias")
    assert db2 > 1e-3


def test_bior_step1_image():
    """Still image, 8x8 patches: step 1 in the bior1.5 basis with the Kaiser
    window, step 2 in the DCT (nldct.cpp:81-148, 640-737)."""
    clean = synth.image(48, 40, seed=100)
    v = np.clip(np.round(synth.add_noise(clean, 20, 200)), 0, 255).astype(np.float32)
    p1 = O.cli_params(1, SIG, 1, px=8, pt=1, wx=9, np_=16)
    p2 = O.cli_params(2, SIG, 1, px=8, pt=1, wx=9, np_=16)
    assert p1.transform == 1 and p2.transform == 0
    B, Bi, _ = O.bases(p1)
    _compare_run(p1, p2, v, nparts=2, bior=(B, Bi))


# ------------------------------------------- predictive search (VBM3D_SEARCH)

def _pred_params(step, F, **kw):
    p = _params(step, F, wx=7, **kw)
    p.sizeSearchWindowPred = 3
    p.nSimilarPatchesPred = 3
    p.dsub = 3.0 if step == 1 else 2.0
    return p


@pytest.fixture
def predictive():
    O.set_predictive(True)
    yield
    O.set_predictive(False)


@pytest.mark.parametrize("nparts", [1, 4])
def test_predictive_search_neighbour_lists(predictive, nparts):
    """The reference's VBM3D_SEARCH block matching (nldct.cpp:904-1130,
    1368-1589): full window in the reference frame, unions of small windows
    around the flow-advanced best matches of the previous frame elsewhere,
    the distance bias at the reference's position, per-frame and final
    std::partial_sort; step 2 sums channel 0 once per channel."""
    v = _int_video(30, 26, 9, c=3, seed=21, hi=127)
    F, _, H, W = v.shape
    ff, bf = _flows(F, H, W, seed=22, push=2.0)
    for step in (1, 2):
        p = _pred_params(step, F)
        idx, nsim, ref, tile = O.bm(p, v, ff, bf, nparts=nparts, tie_mode=0)
        m = M.Mini(p, p, predictive=True)
        groups = m.bm_all(step, v, ff, bf, nparts=nparts)
        assert len(groups) == len(ref) > 50
        bad = [g for g, (ti, r, lst) in enumerate(groups)
               if ti != tile[g] or r != ref[g] or list(idx[g, :nsim[g]]) != lst]
        assert not bad, (step, len(bad))
    # the lists differ from the exhaustive search
    O.set_predictive(False)
    idx2, nsim2, _, _ = O.bm(_pred_params(1, F), v, ff, bf, nparts=nparts)
    O.set_predictive(True)
    idx1, nsim1, _, _ = O.bm(_pred_params(1, F), v, ff, bf, nparts=nparts)
    assert not (np.array_equal(idx1, idx2) and np.array_equal(nsim1, nsim2))


@pytest.mark.parametrize("mistake", ["pred_bias_reference_frame", "pred_all_channels"])
def test_predictive_pins_are_sensitive(predictive, mistake):
    v = _int_video(30, 26, 9, c=3, seed=21, hi=127)
    F, _, H, W = v.shape
    ff, bf = _flows(F, H, W, seed=22, push=2.0)
    step = 2 if mistake == "pred_all_channels" else 1
    p = _pred_params(step, F)
    idx, nsim, ref, tile = O.bm(p, v, ff, bf, nparts=1, tie_mode=0)
    groups = M.Mini(p, p, predictive=True, mistake=mistake).bm_all(step, v, ff, bf, nparts=1)
    assert any(list(idx[g, :nsim[g]]) != lst for g, (_, _, lst) in enumerate(groups))


def test_predictive_two_steps(predictive):
    v = _int_video(30, 26, 9, seed=23)
    F, _, H, W = v.shape
    ff, bf = _flows(F, H, W, seed=24, push=2.0)
    p1, p2 = _pred_params(1, F), _pred_params(2, F)
    ob, of_, orat = O.run(p1, p2, v, ff, bf, nparts=2)
    none = O.cli_params(2, SIG, F, px=0)
    mb, _, r1 = M.Mini(p1, none, predictive=True).run(v, ff, bf, nparts=2)
    _, mf, r2 = M.Mini(O.cli_params(1, SIG, F, px=0), p2, predictive=True).run(v, ff, bf, basic=ob,
                                                                                nparts=2)
    assert _maxdiff(mb, ob) < 1e-3 and _maxdiff(mf, of_) < 1e-3
    assert abs(r1[0] - orat[0]) < 1e-6 * orat[0] and abs(r2[1] - orat[1]) < 1e-6 * orat[1]


def test_predictive_rejects_regions_smaller_than_the_kept_count(predictive):
    v = _int_video(30, 26, 9, seed=23)
    p1 = _pred_params(1, 9)
    p1.nSimilarPatchesPred = 10   # more than a 3x3 predictive window holds
    with pytest.raises(RuntimeError, match="nSimilarPatchesPred"):
        O.run(p1, O.cli_params(2, SIG, 9, px=0), v, nparts=2)
