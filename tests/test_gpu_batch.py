"""Batches of uneven gray sequences (one launch set, replayed from a captured
hipGraph) and the autotuner.  The bar: every sequence's output is bitwise the
single-sequence output (same nparts), on every graph replay, and every
option set the autotuner may pick gives bitwise the same output."""
import numpy as np
import pytest
import torch

from conftest import bm4dof_amd as A, synth

pytestmark = pytest.mark.gpu

SHAPES = [  # (w, h, f, sigma, flow)
    (64, 48, 6, 10.0, False),
    (80, 56, 8, 20.0, True),
    (48, 40, 5, 40.0, False),
    (96, 64, 7, 20.0, True),
    (64, 48, 3, 20.0, False),   # fewer frames than the temporal patch: zeros
    (72, 40, 9, 10.0, True),
]


def _flow(F, H, W, seed, amp=2.0):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    fl = np.zeros((F, 2, H, W), np.float32)
    for t in range(F):
        for c in range(2):
            a, b, ph = rng.uniform(-1, 1, 3)
            fl[t, c] = amp * np.sin(2 * np.pi * (a * xx / W + b * yy / H) + ph)
    return fl


def _make(seed0=0):
    seqs = []
    for i, (w, h, f, sigma, flow) in enumerate(SHAPES):
        clean = synth.video(w, h, f, seed=seed0 + 10 * i)
        noisy = synth.add_noise(clean, sigma, seed0 + 10 * i + 1)
        d = {"shape": noisy.shape, "sigma": sigma,
             "noisy": torch.from_numpy(noisy).cuda(),
             "basic": torch.zeros(noisy.shape, dtype=torch.float32, device="cuda"),
             "final": torch.zeros(noisy.shape, dtype=torch.float32, device="cuda")}
        d["ff"] = torch.from_numpy(_flow(f, h, w, i)).cuda() if flow else None
        d["bf"] = torch.from_numpy(_flow(f, h, w, 100 + i)).cuda() if flow else None
        d["p1"] = A.cli_params(1, sigma, noisy.shape, px=4, pt=4, wx=9, np_=16)
        d["p2"] = A.cli_params(2, sigma, noisy.shape, px=4, pt=4, wx=9, np_=16)
        seqs.append(d)
    return seqs


def _ptr(t):
    return t.data_ptr() if t is not None else None


def _single(ctx, d):
    b = torch.zeros_like(d["noisy"])
    f = torch.zeros_like(d["noisy"])
    ctx.run_dev(d["p1"], d["p2"], d["noisy"].data_ptr(), d["shape"], _ptr(d["ff"]), _ptr(d["bf"]),
                b.data_ptr(), f.data_ptr(), nparts=2)
    torch.cuda.synchronize()
    return b.cpu().numpy(), f.cpu().numpy()


def _seq(d):
    F, _, H, W = d["shape"]
    return A.Seq(d["noisy"].data_ptr(), _ptr(d["ff"]), _ptr(d["bf"]), d["basic"].data_ptr(),
                 d["final"].data_ptr(), W, H, F)


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_single_sequences(ctx, graph):
    seqs = _make()
    refs = [_single(ctx, d) for d in seqs]
    batch = ctx.batch([_seq(d) for d in seqs], [d["p1"] for d in seqs], [d["p2"] for d in seqs],
                      nparts=2, graph=graph)
    try:
        for replay in range(3):
            for d in seqs:
                d["basic"].zero_()
                d["final"].zero_()
            torch.cuda.synchronize()
            ms = batch.run()
            assert ms > 0
            for i, (d, (rb, rf)) in enumerate(zip(seqs, refs)):
                assert np.array_equal(d["basic"].cpu().numpy(), rb), (replay, i)
                assert np.array_equal(d["final"].cpu().numpy(), rf), (replay, i)
            # eager work on the same context between replays
            _single(ctx, seqs[1])
    finally:
        batch.close()
    assert not refs[4][1].any()


def test_graph_batch_after_buffers_grow():
    """A captured batch holds the context buffers' addresses; a run of a
    larger video on the same context between replays makes the context
    grow (move) them.  The next run captures the launch set again and still
    gives the single-sequence outputs.  A fresh context, so that the
    buffers really grow."""
    ctx = A.Context(0)
    try:
        seqs = _make(seed0=500)[:3]
        refs = [_single(ctx, d) for d in seqs]
        batch = ctx.batch([_seq(d) for d in seqs], [d["p1"] for d in seqs], [d["p2"] for d in seqs],
                          nparts=2, graph=True)
        clean = synth.video(160, 120, 10, seed=77)
        big = {"shape": (10, 1, 120, 160), "noisy": torch.from_numpy(synth.add_noise(clean, 20.0, 78)).cuda(),
               "ff": None, "bf": None,
               "p1": A.cli_params(1, 20.0, (10, 1, 120, 160), px=4, pt=4, wx=9, np_=16),
               "p2": A.cli_params(2, 20.0, (10, 1, 120, 160), px=4, pt=4, wx=9, np_=16)}
        try:
            for step in range(3):
                if step == 1:
                    _single(ctx, big)
                for d in seqs:
                    d["basic"].zero_()
                    d["final"].zero_()
                torch.cuda.synchronize()
                batch.run()
                st = batch.stats()
                # replayed as captured; captured again after the growth; then replayed
                assert st["path"] == (2 if step == 1 else 1), (step, st)
                assert st["captures"] == (1 if step == 0 else 2), (step, st)
                for i, (d, (rb, rf)) in enumerate(zip(seqs, refs)):
                    assert np.array_equal(d["basic"].cpu().numpy(), rb), (step, i)
                    assert np.array_equal(d["final"].cpu().numpy(), rf), (step, i)
        finally:
            batch.close()
    finally:
        ctx.close()


def test_batch_with_equal_distances_runs_the_mask_scan(ctx):
    """Integer-valued sequences with constant regions: equal distances in
    heap order put another patch first in some groups, and the reference's
    processing-mask scan skips positions (nldct.cpp:2325-2329).  The launch
    set flags those sequences and the run finishes them through the eager
    path: bitwise equal to single runs and to the CPU oracle."""
    import oracle as O
    seqs = []
    for i, (w, h, f) in enumerate([(56, 40, 7), (64, 48, 6)]):
        clean = synth.video(w, h, f, seed=300 + i)
        noisy = np.clip(np.round(synth.add_noise(clean, 20.0, 301 + i)), 0, 255).astype(np.float32)
        if i == 0:
            noisy[:, :, :20, :24] = 50.0
        d = {"shape": noisy.shape, "np": noisy,
             "noisy": torch.from_numpy(noisy).cuda(),
             "basic": torch.zeros(noisy.shape, dtype=torch.float32, device="cuda"),
             "final": torch.zeros(noisy.shape, dtype=torch.float32, device="cuda"), "ff": None, "bf": None,
             "p1": A.cli_params(1, 20.0, noisy.shape, px=4, pt=4, wx=9, np_=16),
             "p2": A.cli_params(2, 20.0, noisy.shape, px=4, pt=4, wx=9, np_=16)}
        seqs.append(d)
    refs = [_single(ctx, d) for d in seqs]
    o1 = O.cli_params(1, 20.0, 7, px=4, pt=4, wx=9, np_=16)
    o2 = O.cli_params(2, 20.0, 7, px=4, pt=4, wx=9, np_=16)
    ob, of_, _ = O.run(o1, o2, seqs[0]["np"], nparts=2)
    assert np.array_equal(refs[0][0], ob) and np.array_equal(refs[0][1], of_, equal_nan=True)
    batch = ctx.batch([_seq(d) for d in seqs], [d["p1"] for d in seqs], [d["p2"] for d in seqs],
                      nparts=2, graph=True)
    try:
        for replay in range(2):
            for d in seqs:
                d["basic"].zero_()
                d["final"].zero_()
            torch.cuda.synchronize()
            batch.run()
            assert batch.stats()["reruns"] >= 1
            for i, (d, (rb, rf)) in enumerate(zip(seqs, refs)):
                assert np.array_equal(d["basic"].cpu().numpy(), rb), (replay, i)
                assert np.array_equal(d["final"].cpu().numpy(), rf, equal_nan=True), (replay, i)
    finally:
        batch.close()


def test_autotuner_choices_give_equal_output(ctx):
    d = _make(seed0=500)[3]
    outs = []
    try:
        for fbm, fr, cap in ((2, 2, 768), (2, 2, 1024), (2, 2, 1536), (1, 1, 1024), (1, 2, 768),
                             (1, 4, 1536), (0, 2, 1024)):
            ctx.set_option("fast_bm", fbm)
            ctx.set_option("bm_frames", fr)
            ctx.set_option("agg_cap", cap)
            outs.append(_single(ctx, d))
    finally:
        ctx.set_option("fast_bm", 2)
        ctx.set_option("bm_frames", 2)
        ctx.set_option("agg_cap", 1024)
    for b, f in outs[1:]:
        assert np.array_equal(b, outs[0][0]) and np.array_equal(f, outs[0][1])
    rep = ctx.autotune(d["p1"], d["p2"], d["noisy"].data_ptr(), d["shape"], _ptr(d["ff"]), _ptr(d["bf"]),
                       d["basic"].data_ptr(), d["final"].data_ptr())
    assert "chosen" in rep and rep.count("ms") >= 7


def test_autotuned_graph_batch_equals_single_sequences(ctx):
    seqs = _make(seed0=900)
    refs = [_single(ctx, d) for d in seqs]
    batch = ctx.batch([_seq(d) for d in seqs], [d["p1"] for d in seqs], [d["p2"] for d in seqs],
                      nparts=2, graph=True, autotune=True)
    try:
        assert "chosen" in batch.report()
        batch.run()
        for d, (rb, rf) in zip(seqs, refs):
            assert np.array_equal(d["final"].cpu().numpy(), rf)
    finally:
        batch.close()
