"""The command-line tools on the GPU: their outputs equal the library's on the
same inputs and parameters, bit for bit.

  tvl1flow      lib/tvl1flow/main.c: default parameters, nscales capped by
                the image size (main.c:159-163)
  tvl1flow-seq  scripts/tvl1flow-seq.sh: file names, directions, end copies,
                existing files kept
  bm4d-of       src/main_vnlb.cpp through the VideoNLB C++ interface:
                flags, override order, DENO_ONLY with a given basic video,
                -t1/-t2 ignored outside -print-prms (main_vnlb.cpp:233-240)
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import bm4dof_amd as A, synth

pytestmark = pytest.mark.gpu


def _tool(name, *args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    print("tool:", name, *args, flush=True)
    r = subprocess.run([A.TOOLS[name], *map(str, args)], capture_output=True, text=True, timeout=240, env=e)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return r


def _cap(nx, ny, z=0.5, n=100):
    N = np.float32(1 + math.log(math.hypot(nx, ny) / 16.0) / math.log(float(np.float32(1) / np.float32(z))))
    return int(N) if N < n else n


def _frames(tmp_path, F=6, W=48, H=40, C=1, seed=31):
    v = synth.add_noise(synth.video(W, H, F, channels=C, seed=seed), 20, seed + 1)
    for t in range(F):
        A.write_image(tmp_path / f"n_{t:03d}.tif", v[t])
    return v


def test_tvl1flow_tool_equals_library(ctx, tmp_path):
    v = _frames(tmp_path, F=2)
    _tool("tvl1flow", tmp_path / "n_000.tif", tmp_path / "n_001.tif", tmp_path / "o.flo")
    uv = A.read_image(tmp_path / "o.flo")
    assert _cap(48, 40) == 2
    u, w, _ = ctx.tvl1(v[0, 0], v[1, 0], tau=0.25, lam=0.15, theta=0.3, nscales=2, fscale=0,
                       zfactor=0.5, warps=5, eps=0.01)
    assert np.array_equal(uv[0], u) and np.array_equal(uv[1], w)


def test_tvl1flow_seq_tool_equals_library(ctx, tmp_path):
    F = 5
    v = _frames(tmp_path, F=F)
    out = tmp_path / "flow"
    _tool("tvl1flow-seq", tmp_path / "n_%03d.tif", 0, F - 1, out)
    kw = dict(tau=0.25, lam=0.2, theta=0.3, nscales=_cap(48, 40), fscale=1, zfactor=0.5, warps=5, eps=0.01)
    ff, bf = ctx.tvl1_sequence(v[:, 0], **kw)
    for i in range(F):
        assert np.array_equal(A.read_image(out / f"tvl1_{i:03d}_f.flo"), ff[i]), i
        assert np.array_equal(A.read_image(out / f"tvl1_{i:03d}_b.flo"), bf[i]), i
    # one pair alone gives the same flow as the pairs solved together
    u, w, _ = ctx.tvl1(v[2, 0], v[3, 0], **kw)
    assert np.array_equal(ff[2, 0], u) and np.array_equal(ff[2, 1], w)
    # existing files are kept; a missing one is recomputed pair by pair
    keep = (out / "tvl1_001_f.flo").read_bytes()
    os.remove(out / "tvl1_002_f.flo")
    os.remove(out / "tvl1_004_f.flo")
    _tool("tvl1flow-seq", tmp_path / "n_%03d.tif", 0, F - 1, out, "fwd")
    assert (out / "tvl1_001_f.flo").read_bytes() == keep
    assert np.array_equal(A.read_image(out / "tvl1_002_f.flo"), ff[2])
    assert np.array_equal(A.read_image(out / "tvl1_004_f.flo"), ff[3])    # end copy


def _flows(ctx, v):
    return ctx.tvl1_sequence(v[:, 0], tau=0.25, lam=0.2, theta=0.3, nscales=_cap(v.shape[3], v.shape[2]),
                             fscale=1, zfactor=0.5, warps=5, eps=0.01)


@pytest.mark.parametrize("nparts", [2, 4])
def test_bm4d_of_tool_equals_library(ctx, tmp_path, nparts):
    F, W, H = 6, 48, 40
    v = _frames(tmp_path, F=F, W=W, H=H)
    ff, bf = _flows(ctx, v)
    for t in range(F):
        A.write_image(tmp_path / f"f_{t:03d}.flo", ff[t])
        A.write_image(tmp_path / f"b_{t:03d}.flo", bf[t])
    flags = ["-px1", 4, "-pt1", 4, "-px2", 4, "-pt2", 4, "-np1", 32, "-np2", 32, "-wx1", 11, "-wx2", 11]
    _tool("bm4d-of", "-i", tmp_path / "n_%03d.tif", "-f", 0, "-l", F - 1, "-sigma", 20, "-verbose", 0,
          "-fof", tmp_path / "f_%03d.flo", "-bof", tmp_path / "b_%03d.flo",
          "-bsic", tmp_path / "bsic_%03d.tif", "-deno", tmp_path / "deno_%03d.tif", "-t1", 5, *flags,
          env={"BM4DOF_NPARTS": str(nparts)})
    shape = (F, 1, H, W)
    p1 = A.cli_params(1, 20, shape, px=4, pt=4, wx=11, np_=32)
    p2 = A.cli_params(2, 20, shape, px=4, pt=4, wx=11, np_=32)
    basic, final, _ = ctx.run(p1, p2, v, ff, bf, nparts=nparts)
    got_b = np.stack([A.read_image(tmp_path / f"bsic_{t:03d}.tif") for t in range(F)])
    got_f = np.stack([A.read_image(tmp_path / f"deno_{t:03d}.tif") for t in range(F)])
    assert np.array_equal(got_b, basic) and np.array_equal(got_f, final)

    # DENO_ONLY: -px1 0 with the basic video given
    _tool("bm4d-of", "-i", tmp_path / "n_%03d.tif", "-f", 0, "-l", F - 1, "-sigma", 20, "-verbose", 0,
          "-fof", tmp_path / "f_%03d.flo", "-bof", tmp_path / "b_%03d.flo", "-b", tmp_path / "bsic_%03d.tif",
          "-bsic", tmp_path / "b2_%03d.tif", "-deno", tmp_path / "d2_%03d.tif",
          "-px1", 0, "-px2", 4, "-pt2", 4, "-np2", 32, "-wx2", 11, env={"BM4DOF_NPARTS": str(nparts)})
    p1z = A.cli_params(1, 20, shape, px=0, pt=1)
    _, final2, _ = ctx.run(p1z, p2, v, ff, bf, basic=basic, nparts=nparts)
    got_f2 = np.stack([A.read_image(tmp_path / f"d2_{t:03d}.tif") for t in range(F)])
    assert np.array_equal(got_f2, final2) and np.array_equal(got_f2, final)


def test_bm4d_of_tool_color_png(ctx, tmp_path):
    """Colour frames in PNG; outputs quantised as the reference writes them."""
    F, W, H = 4, 32, 24
    v = synth.add_noise(synth.video(W, H, F, channels=3, seed=8), 20, 9)
    q = np.clip(v + 0.5, 0, 255).astype(np.uint8).astype(np.float32)
    for t in range(F):
        A.write_image(tmp_path / f"c_{t:03d}.png", v[t])
    r = _tool("bm4d-of", "-i", tmp_path / "c_%03d.png", "-f", 0, "-l", F - 1, "-sigma", 20,
              "-bsic", tmp_path / "bs_%03d.png", "-deno", tmp_path / "dn_%03d.png", "-px1", 4, "-pt1", 2,
              "-px2", 4, "-pt2", 2, "-np1", 16, "-np2", 16, "-wx1", 7, "-wx2", 7, env={"BM4DOF_NPARTS": "2"})
    assert "Running BM4D-OF on the noisy video" in r.stdout and "Parameters for step 2" in r.stdout
    assert "Processed" in r.stdout and "2nd step." in r.stdout
    shape = (F, 3, H, W)
    p1 = A.cli_params(1, 20, shape, px=4, pt=2, wx=7, np_=16)
    p2 = A.cli_params(2, 20, shape, px=4, pt=2, wx=7, np_=16)
    basic, final, _ = ctx.run(p1, p2, q, nparts=2)
    qb = np.clip(basic + 0.5, 0, 255).astype(np.uint8)
    qf = np.clip(final + 0.5, 0, 255).astype(np.uint8)
    got_b = np.stack([A.read_image(tmp_path / f"bs_{t:03d}.png") for t in range(F)])
    got_f = np.stack([A.read_image(tmp_path / f"dn_{t:03d}.png") for t in range(F)])
    assert np.array_equal(got_b, qb) and np.array_equal(got_f, qf)
