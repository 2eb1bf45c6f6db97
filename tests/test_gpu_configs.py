"""Every benchmark configuration at its own workload, on the GPU.

* 256x256 image, px 8 (bior1.5 step 1), np 16, 21x21 search, beta1 3:
  bitwise equal to the CPU oracle.
* 960x540x20 gray video, sigma 20, px 4, pt 4, np 32, 11x11, +-4 frames along
  TV-L1 flow (5 scales, 10 warps, lambda 0.15): TV-L1 of full-size pairs
  against the oracle; the two-step denoiser at full size (determinism, PSNR)
  and on a 480x270x20 crop with the full-size flows bitwise equal to the
  oracle; end to end (GPU flow and denoiser against the oracle's flow and
  denoiser) within 0.01 dB on a crop.
* 32 uneven sequences (sizes {320x240 .. 1280x720}, 3-30 frames, sigma
  {10, 20, 40}) in one captured graph with the autotuner: every sequence
  finite with a PSNR gain, a sample bitwise equal to single runs.
* learned single pass at 960x540x20: finite, PSNR against the plain path
  reported, gain over the first step.
* colour 1920x1080x30, sigma 40, np 48, 15x15, +-6 frames: plain, learned
  and learned with fallback, by properties (colour video has no CPU parity,
  the reference's mask clear is undefined there).

Each test prints what it runs before each GPU step.
"""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

import oracle as O
from conftest import bm4dof_amd as A, synth
from test_gpu_denoise import _same

pytestmark = pytest.mark.gpu

W1, H1, F1 = 960, 540, 20
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _say(*a):
    print("[configs]", *a, flush=True)


def _psnr(a, b):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return float(10 * np.log10(255.0 ** 2 / np.mean(d * d)))


# ------------------------------------------------------- 256x256 image

def test_image_256_bior_parity(ctx):
    clean = synth.image(256, 256, seed=100)
    noisy = synth.add_noise(clean, 20, 200)
    kw1 = dict(px=8, pt=1, wx=21, np_=16, beta=3.0)
    kw2 = dict(px=8, pt=1, wx=21, np_=16)
    a1, a2 = A.cli_params(1, 20, noisy.shape, **kw1), A.cli_params(2, 20, noisy.shape, **kw2)
    o1, o2 = O.cli_params(1, 20, 1, **kw1), O.cli_params(2, 20, 1, **kw2)
    assert a1.transform == 1 and a1.sizeSearchWindow == 21 and a1.nSimilarPatches == 16
    _say("image 256x256: GPU two steps, nparts 2")
    gb, gf, ggr = ctx.run(a1, a2, noisy, nparts=2)
    ob, of_, ogr = O.run(o1, o2, noisy, nparts=2)
    _same(gb, ob, "basic")
    _same(gf, of_, "final")
    assert ggr == pytest.approx(ogr, abs=1e-5)
    assert _psnr(gf, clean) > _psnr(noisy, clean) + 7


# ------------------------------------------------ 960x540x20 gray video

@pytest.fixture(scope="module")
def video1(ctx):
    t = time.perf_counter()
    clean = synth.video(W1, H1, F1, seed=101, threads=THREADS)
    noisy = synth.add_noise(clean, 20, 201)
    _say(f"960x540x20 video generated in {time.perf_counter() - t:.1f} s; GPU TV-L1 of 38 pairs")
    ff, bf = ctx.tvl1_sequence(noisy[:, 0], lam=0.15, nscales=5, warps=10)
    return clean, noisy, ff, bf


def _p(step, shape, sigma=20.0, **kw):
    kw = dict(dict(px=4, pt=4, wx=11, np_=32), **kw)
    return A.cli_params(step, sigma, shape, **kw), O.cli_params(step, sigma, shape[0], **kw)


def test_tvl1_full_size_pairs(ctx, video1):
    """Two 960x540 pairs (forward and backward of frames 9, 10) with the
    workload's parameters against the oracle; the sequence solve gives the
    same flows as the single-pair solve."""
    _, noisy, ff, bf = video1
    kw = dict(lam=0.15, nscales=5, warps=10)
    pairs = [(9, 10), (10, 9)]
    _say("TV-L1 960x540 pairs on the GPU")
    gpu = [ctx.tvl1(noisy[i, 0], noisy[j, 0], **kw) for i, j in pairs]
    with ThreadPoolExecutor(2) as ex:   # the oracle's TV-L1 is single-threaded
        cpu = list(ex.map(lambda ij: O.tvl1(noisy[ij[0], 0], noisy[ij[1], 0], **kw), pairs))
    for (i, j), (u, v, it), (uo, vo, ito) in zip(pairs, gpu, cpu):
        assert np.abs(it.astype(int) - ito).max() <= 1, (i, j)
        epe = np.sqrt((u - uo) ** 2 + (v - vo) ** 2)
        if np.array_equal(it, ito):
            assert np.array_equal(u, uo) and np.array_equal(v, vo), (i, j, epe.max())
        else:
            assert epe.mean() < 1e-3, (i, j, epe.mean())
        seq = ff[i] if j == i + 1 else bf[i]
        assert np.sqrt((seq[0] - u) ** 2 + (seq[1] - v) ** 2).mean() < 1e-3, (i, j)


def test_full_size_two_steps_with_flow(ctx, video1):
    clean, noisy, ff, bf = video1
    a1, _ = _p(1, noisy.shape)
    a2, _ = _p(2, noisy.shape)
    _say("960x540x20 two steps with TV-L1 flow, twice")
    b1, f1, gr = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    b2, f2, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    assert np.array_equal(b1, b2) and np.array_equal(f1, f2)
    assert np.isfinite(f1).all() and np.isfinite(b1).all()
    pn, pb, pf = _psnr(noisy, clean), _psnr(b1, clean), _psnr(f1, clean)
    _say(f"PSNR noisy {pn:.3f} basic {pb:.3f} final {pf:.3f}, groups ratio {gr}")
    assert pf > pn + 17 and pf > pb + 1.5
    # with flow the trajectories follow the motion: better than without
    _, f0, _ = ctx.run(a1, a2, noisy, nparts=2)
    assert pf > _psnr(f0, clean)


def test_crop_with_full_size_flows_equals_oracle(ctx, video1):
    """480x270x20 centre crop of the workload and of its TV-L1 flows, nparts
    2 x threads as the reference on that many threads: bitwise equal to the
    oracle."""
    clean, noisy, ff, bf = video1
    cw, ch = 480, 270
    x0, y0 = (W1 - cw) // 2, (H1 - ch) // 2
    sl = (slice(None), slice(None), slice(y0, y0 + ch), slice(x0, x0 + cw))
    n, f1, b1 = (np.ascontiguousarray(a[sl]) for a in (noisy, ff, bf))
    a1, o1 = _p(1, n.shape)
    a2, o2 = _p(2, n.shape)
    nparts = 2 * THREADS
    _say(f"crop 480x270x20, nparts {nparts}: GPU, then oracle on {THREADS} threads")
    gb, gf, ggr = ctx.run(a1, a2, n, f1, b1, nparts=nparts)
    prev = O._omp_threads(THREADS)
    try:
        ob, of_, ogr = O.run(o1, o2, n, f1, b1, nparts=nparts)
    finally:
        O._omp_threads(prev)
    _same(gb, ob, "basic")
    _same(gf, of_, "final")
    assert ggr == pytest.approx(ogr, abs=1e-4)


def _oracle_flows(frames, **kw):
    """The oracle's TV-L1 of every pair of a sequence with the end copies of
    scripts/tvl1flow-seq.sh:31-61, pairs on a thread pool."""
    F, H, W = frames.shape
    ff = np.zeros((F, 2, H, W), np.float32)
    bf = np.zeros((F, 2, H, W), np.float32)
    jobs = [(i, i + 1) for i in range(F - 1)] + [(i, i - 1) for i in range(1, F)]
    with ThreadPoolExecutor(THREADS) as ex:   # the oracle's TV-L1 is single-threaded
        res = list(ex.map(lambda ij: O.tvl1(frames[ij[0]], frames[ij[1]], **kw), jobs))
    for (i, j), (u, v, _) in zip(jobs, res):
        dst = ff if j == i + 1 else bf
        dst[i, 0], dst[i, 1] = u, v
    ff[F - 1] = ff[F - 2]
    bf[0] = bf[1]
    return ff, bf


def test_end_to_end_flow_and_denoiser_within_001_db(ctx, video1):
    """GPU flow -> GPU denoiser against CPU flow -> CPU denoiser (the oracle
    for both) on a 256x144x10 crop of the workload: PSNR within 0.01 dB.
    The flows are not bit-gated (the stopping test's reduction order), so
    max-abs is reported, not gated."""
    clean, noisy, _, _ = video1
    cw, ch, cf = 256, 144, 10
    x0, y0 = 352, 198
    sl = (slice(0, cf), slice(None), slice(y0, y0 + ch), slice(x0, x0 + cw))
    n, c = np.ascontiguousarray(noisy[sl]), clean[sl]
    kw = dict(lam=0.15, nscales=5, warps=10)
    _say("end to end 256x144x10: GPU flows and denoiser")
    gff, gbf = ctx.tvl1_sequence(n[:, 0], **kw)
    a1, o1 = _p(1, n.shape)
    a2, o2 = _p(2, n.shape)
    _, gf, _ = ctx.run(a1, a2, n, gff, gbf, nparts=2)
    _say("end to end: oracle flows of 18 pairs and oracle denoiser")
    off, obf = _oracle_flows(n[:, 0], **kw)
    _, of_, _ = O.run(o1, o2, n, off, obf, nparts=2)
    dpsnr = _psnr(gf, c) - _psnr(of_, c)
    epe = np.sqrt(((gff - off) ** 2).sum(axis=1)).mean()
    _say(f"end to end: dPSNR {dpsnr:+.5f} dB, max-abs {np.abs(gf - of_).max():.4f}, "
         f"mean flow EPE {epe:.2e}")
    assert abs(dpsnr) <= 0.01


# ------------------------------------------------ 32 uneven sequences

def test_batch_32_uneven_sequences(ctx):
    rng = np.random.default_rng(2024)
    sizes = [(320, 240), (640, 360), (960, 540), (1280, 720)]
    seqs, keep, p1s, p2s, info = [], [], [], [], []
    t = time.perf_counter()
    for i in range(32):
        w, h = sizes[rng.integers(0, 4)]
        f = int(rng.integers(3, 31))
        sigma = float(rng.choice([10.0, 20.0, 40.0]))
        clean = synth.video(w, h, f, seed=7000 + i, threads=THREADS)
        noisy = synth.add_noise(clean, sigma, 8000 + i)
        dn = torch.from_numpy(noisy).cuda()
        db, df = torch.full_like(dn, -1.0), torch.full_like(dn, -1.0)
        keep.append((dn, db, df))
        seqs.append(A.Seq(dn.data_ptr(), None, None, db.data_ptr(), df.data_ptr(), w, h, f))
        p1s.append(A.cli_params(1, sigma, noisy.shape, px=4, pt=4, wx=11, np_=32))
        p2s.append(A.cli_params(2, sigma, noisy.shape, px=4, pt=4, wx=11, np_=32))
        info.append((w, h, f, sigma, clean, noisy))
    _say(f"32 sequences generated in {time.perf_counter() - t:.1f} s; capture with autotune")
    assert {(w, h) for w, h, *_ in info} == set(sizes)
    batch = ctx.batch(seqs, p1s, p2s, nparts=2, graph=True, autotune=True)
    try:
        _say("batch replays")
        ms = [batch.run() for _ in range(2)]
        torch.cuda.synchronize()
        _say(f"replays {ms} ms, {sum(w * h * f for w, h, f, *_ in info) / 1e6:.1f} Mpix")
    finally:
        batch.close()
    for i, (w, h, f, sigma, clean, noisy) in enumerate(info):
        fin = keep[i][2].cpu().numpy()
        if f < 4:   # fewer frames than the temporal patch: no group, zeros
            assert not fin.any(), i
            continue
        assert np.isfinite(fin).all(), i
        assert _psnr(fin, clean) > _psnr(noisy, clean) + 5, (i, w, h, f, sigma)
    # a sample against single-sequence runs: the largest, the smallest and
    # two others, bitwise
    order = sorted(range(32), key=lambda i: info[i][0] * info[i][1] * info[i][2])
    for i in {order[0], order[-1], order[10], order[21]}:
        w, h, f, sigma, _, noisy = info[i]
        dn, db, df = keep[i]
        sb, sf = torch.empty_like(dn), torch.empty_like(dn)
        _say(f"single run of sequence {i}: {w}x{h}x{f} sigma {sigma:g}")
        ctx.run_dev(p1s[i], p2s[i], dn.data_ptr(), noisy.shape, None, None, sb.data_ptr(),
                    sf.data_ptr(), nparts=2)
        torch.cuda.synchronize()
        assert torch.equal(sb, db) and torch.equal(sf, df), i


# ------------------------------------------------- learned single pass

def test_learned_pass_full_size(ctx, video1):
    from conftest import _load
    L = _load("bm4dof_amd.learned", os.path.join(os.path.dirname(A.__file__), "learned.py"))
    clean, noisy, ff, bf = video1
    a1, _ = _p(1, noisy.shape)
    a2, _ = _p(2, noisy.shape)
    _say("learned single pass 960x540x20 and plain two steps")
    ctx.mlp_set(L.load())
    out, _ = ctx.run_learned(a1, a2, noisy, ff, bf, nparts=2)
    basic, final, _ = ctx.run(a1, a2, noisy, ff, bf, nparts=2)
    assert np.isfinite(out).all()
    pl, pp, pb = _psnr(out, clean), _psnr(final, clean), _psnr(basic, clean)
    _say(f"learned {pl:.3f} dB, plain two-step {pp:.3f} dB (delta {pl - pp:+.3f}), "
         f"plain step 1 {pb:.3f} dB")
    assert pl > _psnr(noisy, clean) + 15
    assert pl > pb


# ------------------------------------------------------ colour 1080p

def test_colour_1080p_plain_learned_fallback(ctx):
    from conftest import _load
    L = _load("bm4dof_amd.learned", os.path.join(os.path.dirname(A.__file__), "learned.py"))
    Wc, Hc, Fc, sig = 1920, 1080, 30, 40.0
    t = time.perf_counter()
    clean = synth.video(Wc, Hc, Fc, channels=3, n_shapes=40, seed=4242, threads=THREADS)
    noisy = synth.add_noise(clean, sig, 4243)
    _say(f"colour 1920x1080x30 generated in {time.perf_counter() - t:.1f} s; luma TV-L1")
    dev = torch.device("cuda:0")
    dn = torch.from_numpy(noisy).to(dev)
    luma = (0.299 * dn[:, 0] + 0.587 * dn[:, 1] + 0.114 * dn[:, 2]).contiguous()
    ff = torch.empty((Fc, 2, Hc, Wc), dtype=torch.float32, device=dev)
    bf = torch.empty_like(ff)
    ctx.tvl1_sequence_dev(luma.data_ptr(), Wc, Hc, Fc, ff.data_ptr(), bf.data_ptr())
    p1 = A.cli_params(1, sig, noisy.shape, px=4, pt=4, wx=15, np_=48)
    p2 = A.cli_params(2, sig, noisy.shape, px=4, pt=4, wx=15, np_=48)
    for p in (p1, p2):
        p.sizeSearchTimeRangeFwd = p.sizeSearchTimeRangeBwd = 6
    assert p1.nSimilarPatches == 48 and p1.colorSpace == 1
    db, dfin, dl = torch.empty_like(dn), torch.empty_like(dn), torch.empty_like(dn)
    _say("colour plain two steps")
    gr = ctx.run_dev(p1, p2, dn.data_ptr(), noisy.shape, ff.data_ptr(), bf.data_ptr(),
                     db.data_ptr(), dfin.data_ptr(), nparts=2)
    plain = dfin.cpu().numpy()
    assert np.isfinite(plain).all()
    pn, pp = _psnr(noisy, clean), _psnr(plain, clean)
    assert pp > pn + 14, (pn, pp)
    # np 48 acts as 32: the power-of-two cut (nldct.cpp:1285-1289)
    assert ctx.stats()["rows2"] <= 32 * ctx.stats()["groups2"]
    ctx.mlp_set(L.load())
    _say("colour learned pass, then with fallback")
    ctx.run_learned_dev(p1, p2, dn.data_ptr(), noisy.shape, ff.data_ptr(), bf.data_ptr(), dl.data_ptr(),
                        nparts=2)
    learned = dl.cpu().numpy()
    frac = ctx.run_learned_dev(p1, p2, dn.data_ptr(), noisy.shape, ff.data_ptr(), bf.data_ptr(),
                               dl.data_ptr(), nparts=2, gvar_thr=0.02)
    fallback = dl.cpu().numpy()
    pl, pf = _psnr(learned, clean), _psnr(fallback, clean)
    _say(f"colour PSNR noisy {pn:.3f} plain {pp:.3f} learned {pl:.3f} fallback {pf:.3f} "
         f"(flagged {frac:.4f}), groups ratio {gr}")
    assert np.isfinite(learned).all() and np.isfinite(fallback).all()
    assert pl > pn + 10 and 0.0 < frac < 1.0 and pf >= pl
