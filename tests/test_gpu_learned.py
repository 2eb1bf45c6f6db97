"""GPU checks of the learned single pass against the CPU float64
restatement of the network (oracle/mlp_oracle.py): matrix-core inference,
training gradients (relative error <= 1e-3), Adam, training on pairs from
the plain two-step method, what the trained network gains over the pass
without it, and the plain two-step fallback."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import REPO, bm4dof_amd as A, synth
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _learned_module():
    spec = importlib.util.spec_from_file_location("bm4dof_amd.learned",
                                                  os.path.join(REPO, "bm4d-of_amd", "learned.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _batch(B, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1.5, (B, 66)).astype(np.float32)
    x[:, 64] = np.log(rng.choice([10, 20, 30, 40], B) / 20.0)
    x[:, 65] = rng.uniform(0, 2, B)
    t = rng.uniform(0, 1, (B, 64)).astype(np.float32)
    return x, t


def test_matrix_core_inference_matches_oracle(ctx):
    theta = M.init_params(11)
    ctx.mlp_set(theta)
    x, _ = _batch(100, 1)          # 100 rows: 4 waves, the last one partial
    g = ctx.mlp_forward(x)
    ref, _ = M.forward(theta.astype(np.float64), x)
    d = np.abs(g - ref)
    # float32 accumulation against float64: an intermediate may round to the
    # neighbouring bf16 value, which moves a gain by ~1e-3 at most
    assert d.max() < 4e-3 and d.mean() < 2e-4, (d.max(), d.mean())
    assert np.array_equal(ctx.mlp_get(), theta)


def test_training_gradient_matches_oracle(ctx):
    theta = M.init_params(12)
    ctx.mlp_set(theta)
    for B, seed in ((64, 2), (4096, 3)):
        x, t = _batch(B, seed)
        loss, g = ctx.mlp_grad(x, t)
        lref, gref = M.backward(theta.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
        rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
        assert rel <= 1e-3, rel
        assert abs(loss - lref) <= 1e-5 * lref


def test_adam_step_matches_oracle(ctx):
    import torch
    theta = M.init_params(13)
    ctx.mlp_set(theta)
    x, t = _batch(256, 4)
    _, g = ctx.mlp_grad(x, t)
    dx = torch.from_numpy(x).cuda()
    dt = torch.from_numpy(t).cuda()
    lr = 1e-3
    ctx.mlp_train_dev(dx.data_ptr(), dt.data_ptr(), 256, 256, 1, lr)   # whole set, in order
    got = ctx.mlp_get().astype(np.float64)
    want, _, _ = M.adam(theta.astype(np.float64), np.zeros(M.N_PARAMS), np.zeros(M.N_PARAMS),
                        g.astype(np.float64), 1, lr)
    assert np.abs(got - want).max() <= 1e-6 + 1e-5 * lr


@pytest.fixture(scope="module")
def trained(ctx):
    """A network trained on pairs from the plain two-step method over four
    small seeded videos at sigma 10..40 (the training script's own code path)."""
    import torch
    L = _learned_module()
    xs, ts = [], []
    for i, sigma in enumerate((10.0, 20.0, 30.0, 40.0)):
        clean = synth.video(96, 80, 10, seed=300 + i)
        noisy = synth.add_noise(clean, sigma, 400 + i)
        x, t = L.training_pairs(ctx, A, noisy, sigma, stride=1)
        xs.append(x)
        ts.append(t)
    x, t = torch.cat(xs), torch.cat(ts)
    losses = L.train(ctx, x, t, steps=3000, batch=4096, lr=2e-3, seed=7)
    return L, ctx.mlp_get(), losses, x.shape[0]


def test_training_pairs_features_and_targets(trained, ctx):
    """Rows of the pairs kernel obey the feature definition
    (oracle/mlp_oracle.features): the row statistic is log1p of the mean
    square of the row's 64 coefficient features, feature 64 is log(sigma/20),
    and targets are gains in [0, 1]."""
    L = trained[0]
    clean = synth.video(64, 48, 6, seed=5)
    noisy = synth.add_noise(clean, 30.0, 6)
    x, t = L.training_pairs(ctx, A, noisy, 30.0, stride=3)
    x, t = x.cpu().numpy().astype(np.float64), t.cpu().numpy()
    assert x.shape[0] > 100 and x.shape[0] == t.shape[0]
    np.testing.assert_allclose(x[:, 64], np.log(30.0 / 20.0), rtol=1e-6)
    np.testing.assert_allclose(x[:, 65], np.log1p((x[:, :64] ** 2).sum(1) / 64.0), rtol=1e-4, atol=1e-5)
    assert t.min() >= 0.0 and t.max() <= 1.0


def test_training_lowers_the_loss(trained):
    _, _, losses, rows = trained
    assert rows > 20000
    assert losses[-100:].mean() < 0.6 * losses[:20].mean()


def test_learned_pass_gains_over_the_pass_without_it(ctx, trained):
    """Held-out video: the single pass with the network against the same
    single pass without it (the first step's hard threshold), and against
    the plain two-step path."""
    _, theta, _, _ = trained
    ctx.mlp_set(theta)
    clean = synth.video(96, 80, 10, seed=999)
    noisy = synth.add_noise(clean, 20.0, 1999)
    p1 = A.cli_params(1, 20, noisy.shape, px=4, pt=4, wx=11, np_=32)
    p2 = A.cli_params(2, 20, noisy.shape, px=4, pt=4, wx=11, np_=32)
    out, frac = ctx.run_learned(p1, p2, noisy)
    basic, final, _ = ctx.run(p1, p2, noisy)
    ps = lambda v: 10 * np.log10(255.0 ** 2 / np.mean((v.astype(np.float64) - clean) ** 2))
    print("psnr noisy %.2f step1-only %.2f learned %.2f plain two-step %.2f" %
          (ps(noisy), ps(basic), ps(out), ps(final)))
    assert np.isfinite(out).all() and frac == 0.0
    assert ps(out) > ps(basic) + 0.1


def test_fallback_takes_the_plain_result(ctx, trained):
    _, theta, _, _ = trained
    ctx.mlp_set(theta)
    clean = synth.video(64, 48, 8, seed=77)
    noisy = synth.add_noise(clean, 20.0, 78)
    p1 = A.cli_params(1, 20, noisy.shape, px=4, pt=4, wx=9, np_=16)
    p2 = A.cli_params(2, 20, noisy.shape, px=4, pt=4, wx=9, np_=16)
    plain = ctx.run(p1, p2, noisy)[1]
    learned, f0 = ctx.run_learned(p1, p2, noisy, gvar_thr=-1.0)
    none, fn = ctx.run_learned(p1, p2, noisy, gvar_thr=1e9)
    assert fn == 0.0 and np.array_equal(none, learned)
    allf, fa = ctx.run_learned(p1, p2, noisy, gvar_thr=0.0)
    assert fa == 1.0 and np.array_equal(allf, plain)
    some, fs = ctx.run_learned(p1, p2, noisy, gvar_thr=0.02)
    assert 0.0 <= fs <= 1.0
    mixed = (some == plain) | (some == learned)
    assert mixed.all()


def test_colour_channels_follow_the_first_channel(ctx, trained):
    """RGB video whose three channels are one gray video: matching and
    weights come from the first channel and every channel is filtered with
    them, so each output channel equals the one-channel learned pass
    bitwise."""
    _, theta, _, _ = trained
    ctx.mlp_set(theta)
    clean = synth.video(64, 48, 6, seed=41)
    gray = synth.add_noise(clean, 20.0, 42)
    rgb = np.repeat(gray, 3, axis=1)
    kw = dict(px=4, pt=4, wx=9, np_=16)
    g1, g2 = A.cli_params(1, 20, gray.shape, **kw), A.cli_params(2, 20, gray.shape, **kw)
    c1, c2 = A.cli_params(1, 20, rgb.shape, **kw), A.cli_params(2, 20, rgb.shape, **kw)
    c1.colorSpace = c2.colorSpace = 0   # RGB: no transform
    want, _ = ctx.run_learned(g1, g2, gray)
    got, frac = ctx.run_learned(c1, c2, rgb)
    assert frac == 0.0
    for c in range(3):
        assert np.array_equal(got[:, c], want[:, 0]), c


@pytest.mark.parametrize("flow", [False, True])
def test_colour_fallback_takes_the_plain_result(ctx, trained, flow):
    """Colour video (YUV, 48 similar patches acting as 32, temporal range
    along the flow): no flagged group gives the learned pass, all flagged
    gives the plain colour two-step result bitwise, and a threshold between
    gives each sample one of the two."""
    _, theta, _, _ = trained
    ctx.mlp_set(theta)
    clean = synth.video(64, 48, 6, channels=3, seed=90)
    noisy = synth.add_noise(clean, 40.0, 91)
    kw = dict(px=4, pt=4, wx=9, np_=48)
    p1, p2 = A.cli_params(1, 40, noisy.shape, **kw), A.cli_params(2, 40, noisy.shape, **kw)
    ff = bf = None
    if flow:
        ff, bf = ctx.tvl1_sequence(noisy[:, 0], nscales=3, warps=3)
        for p in (p1, p2):
            p.sizeSearchTimeRangeFwd = p.sizeSearchTimeRangeBwd = 2
    plain = ctx.run(p1, p2, noisy, ff, bf)[1]
    learned, f0 = ctx.run_learned(p1, p2, noisy, ff, bf, gvar_thr=-1.0)
    none, fn = ctx.run_learned(p1, p2, noisy, ff, bf, gvar_thr=1e9)
    assert f0 == 0.0 and fn == 0.0 and np.array_equal(none, learned)
    allf, fa = ctx.run_learned(p1, p2, noisy, ff, bf, gvar_thr=0.0)
    assert fa == 1.0 and np.array_equal(allf, plain)
    some, fs = ctx.run_learned(p1, p2, noisy, ff, bf, gvar_thr=0.02)
    assert 0.0 < fs < 1.0
    assert ((some == plain) | (some == learned)).all()
    ps = lambda v: 10 * np.log10(255.0 ** 2 / np.mean((v.astype(np.float64) - clean) ** 2))
    print("colour psnr noisy %.2f learned %.2f plain %.2f blend %.2f (flagged %.3f)" %
          (ps(noisy), ps(learned), ps(plain), ps(some), fs))
    assert ps(learned) > ps(noisy) + 5.0


def test_held_out_evaluation_reports_both_paths(ctx, trained):
    """learned.evaluate: PSNR of the plain two-step path and of the learned
    pass on held-out seeded sequences, per noise level, with device times."""
    L, theta, _, _ = trained
    res, summary = L.evaluate(ctx, A, synth, theta, n=2, size=(96, 80, 8), sigmas=(20.0, 40.0))
    assert len(res) == 2 and summary["sequences"] == 2
    for r in res:
        assert r["psnr_plain"] > r["psnr_noisy"] + 3 and r["psnr_learned"] > r["psnr_noisy"] + 3
        assert r["ms_plain"] > 0 and r["ms_learned"] > 0
    assert set(summary["dpsnr_by_sigma_db"]) == {"20.0", "40.0"}
    assert np.isfinite(summary["dpsnr_mean_db"])


def test_learned_pass_with_8x8x1_patches(ctx, trained):
    """The learned kernel's 8 x 8 x 1 form (one frame per patch): finite,
    well above the noisy input, and with every group flagged it gives the
    plain two-step result bitwise."""
    _, theta, _, _ = trained
    ctx.mlp_set(theta)
    clean = synth.video(72, 56, 4, seed=404)
    noisy = synth.add_noise(clean, 25.0, 405)
    p1 = A.cli_params(1, 25, noisy.shape, px=8, pt=1, wx=15, np_=16)
    p2 = A.cli_params(2, 25, noisy.shape, px=8, pt=1, wx=15, np_=16)
    out, frac = ctx.run_learned(p1, p2, noisy)
    ps = lambda v: 10 * np.log10(255.0 ** 2 / np.mean((v.astype(np.float64) - clean) ** 2))
    assert np.isfinite(out).all() and frac == 0.0
    assert ps(out) > ps(noisy) + 5.0
    allf, fa = ctx.run_learned(p1, p2, noisy, gvar_thr=0.0)
    assert fa == 1.0 and np.array_equal(allf, ctx.run(p1, p2, noisy)[1])
