"""The allocation-free flow entry point, bm4dof_tvl1_noalloc, through the
C-ABI: bitwise the flows and iteration counts of bm4dof_tvl1 on device
buffers, on two calls in a row (the second on a smaller pair, so that it
runs on what the first left in the shared scratch), and a clear error for a
pair whose scratch does not fit."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import bm4dof_amd as A, synth

pytestmark = pytest.mark.gpu


def _pair(w, h, seed):
    v = synth.add_noise(synth.video(w, h, 2, seed=seed), 20, seed + 1)
    return v[0, 0].copy(), v[1, 0].copy()


def _noalloc(ctx, I0, I1, prm):
    H, W = I0.shape
    d0 = torch.from_numpy(I0).cuda()
    d1 = torch.from_numpy(I1).cuda()
    u = torch.full((H, W), 7.0, dtype=torch.float32, device="cuda")
    v = torch.full((H, W), 7.0, dtype=torch.float32, device="cuda")
    it = np.zeros(prm.nscales * prm.warps, np.int32)
    torch.cuda.synchronize()
    rc = A.lib().bm4dof_tvl1_noalloc(ctx.h, C.byref(prm), d0.data_ptr(), d1.data_ptr(), u.data_ptr(),
                                     v.data_ptr(), W, H, it.ctypes.data)
    ctx._check(rc)
    torch.cuda.synchronize()
    return u.cpu().numpy(), v.cpu().numpy(), it.reshape(prm.nscales, prm.warps)


@pytest.mark.parametrize("order", ["large_then_small", "small_then_large"])
def test_noalloc_flow_equals_the_allocating_entry_point(order):
    ctx = A.Context(0)
    try:
        prm = A.TVL1Params(0.25, 0.15, 0.3, 5, 0, 0.5, 10, 0.01, 0)
        sizes = [(960, 540, 11), (320, 240, 12)]
        if order == "small_then_large":
            sizes = sizes[::-1]
        for w, h, seed in sizes:
            I0, I1 = _pair(w, h, seed)
            print(f"no-alloc TV-L1 {w}x{h}", flush=True)
            u, v, it = _noalloc(ctx, I0, I1, prm)
            ur, vr, itr = ctx.tvl1(I0, I1, lam=0.15, nscales=5, warps=10)
            assert np.array_equal(it, itr), (w, h)
            assert np.array_equal(u, ur) and np.array_equal(v, vr), (w, h)
            assert np.abs(u).max() > 0
    finally:
        ctx.close()


def test_noalloc_rejects_pairs_beyond_the_scratch():
    ctx = A.Context(0)
    try:
        prm = A.TVL1Params(0.25, 0.15, 0.3, 5, 0, 0.5, 10, 0.01, 0)
        w, h = 1280, 720
        d = torch.zeros((h, w), dtype=torch.float32, device="cuda")
        u = torch.zeros_like(d)
        rc = A.lib().bm4dof_tvl1_noalloc(ctx.h, C.byref(prm), d.data_ptr(), d.data_ptr(), u.data_ptr(),
                                         u.data_ptr(), w, h, None)
        assert rc == -1
        msg = A.lib().bm4dof_last_error(ctx.h).decode()
        assert "needs" in msg and str(48 << 20) in msg, msg
    finally:
        ctx.close()
