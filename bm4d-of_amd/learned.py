"""Training driver of the learned single pass.

The network (66 -> 128 -> 128 -> 64, ReLU, sigmoid, bf16 weights) maps the
Haar-domain step-1 coefficients of every group row to the step-2 Wiener
gains of the plain two-step method (src/core/nldct.cpp:2131), so that one
pass (step 1's grid and matching, the network's gains, step 2's weights)
replaces the two.  Everything that computes runs in libbm4dof.so:

  * training pairs: the plain two-step path on a noisy training video gives
    its basic estimate; bm4dof_learned_pairs then takes, at the neighbours
    of every stride-th group of the first step's grid, the noisy
    coefficients (features) and the basic estimate's Wiener gains (targets);
  * training: bm4dof_mlp_train, Adam on minibatches (HIP kernels);
  * inference: the network runs inside the group kernel
    (bm4dof_run_learned), on the bf16 matrix cores.

This module only drives those calls; torch provides device memory.

    python -m bm4dof_amd.learned --out weights.f32     (see --help)
"""
import argparse
import os
import time

import numpy as np

N_IN, N_H, N_OUT = 66, 128, 64
N_PARAMS = N_IN * N_H + N_H + N_H * N_H + N_H + N_H * N_OUT + N_OUT
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "gain_mlp_66_128_128_64.f32")
MAGIC = b"BM4DOFMLP1\0\0"


def init_theta(seed=0):
    """He initialisation, biases zero; W[k][j] input-major."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / N_IN), (N_IN, N_H)), np.zeros(N_H),
             rng.normal(0, np.sqrt(2.0 / N_H), (N_H, N_H)), np.zeros(N_H),
             rng.normal(0, np.sqrt(1.0 / N_H), (N_H, N_OUT)), np.zeros(N_OUT)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def save(theta, path):
    theta = np.ascontiguousarray(theta, np.float32)
    assert theta.size == N_PARAMS
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.array([N_IN, N_H, N_OUT], np.int32).tobytes())
        f.write(theta.tobytes())


def load(path=DEFAULT_WEIGHTS):
    with open(path, "rb") as f:
        if f.read(len(MAGIC)) != MAGIC:
            raise ValueError(f"{path}: not a gain-network weight file")
        dims = np.frombuffer(f.read(12), np.int32)
        if tuple(dims) != (N_IN, N_H, N_OUT):
            raise ValueError(f"{path}: network shape {tuple(dims)}")
        theta = np.frombuffer(f.read(4 * N_PARAMS), np.float32).copy()
    if theta.size != N_PARAMS:
        raise ValueError(f"{path}: truncated")
    return theta


def training_pairs(ctx, A, noisy, sigma, px=4, pt=4, wx=11, np_=32, stride=8, nparts=2, fflow=None,
                   bflow=None):
    """(x, t) device tensors of the valid rows of one noisy video (F, 1, H, W)."""
    import torch
    F, C, H, W = noisy.shape
    dev = torch.device("cuda", ctx.device)
    d_noisy = torch.from_numpy(np.ascontiguousarray(noisy, np.float32)).to(dev)
    d_basic = torch.empty_like(d_noisy)
    d_final = torch.empty_like(d_noisy)
    p1 = A.cli_params(1, sigma, noisy.shape, px=px, pt=pt, wx=wx, np_=np_)
    p2 = A.cli_params(2, sigma, noisy.shape, px=px, pt=pt, wx=wx, np_=np_)
    dff = dbf = None
    if fflow is not None:
        dff = torch.from_numpy(np.ascontiguousarray(fflow, np.float32)).to(dev)
        dbf = torch.from_numpy(np.ascontiguousarray(bflow, np.float32)).to(dev)
    fp = dff.data_ptr() if dff is not None else None
    bp = dbf.data_ptr() if dbf is not None else None
    ctx.run_dev(p1, p2, d_noisy.data_ptr(), noisy.shape, fp, bp, d_basic.data_ptr(), d_final.data_ptr(),
                nparts=nparts)
    # upper bound of the sampled rows: the grid has at most W*H*F positions
    max_rows = ((W * H * F) // stride + 1) * 32
    x = torch.empty((max_rows, N_IN), dtype=torch.float32, device=dev)
    t = torch.empty((max_rows, N_OUT), dtype=torch.float32, device=dev)
    fl = torch.empty((max_rows,), dtype=torch.uint8, device=dev)
    rows = ctx.learned_pairs_dev(p1, d_noisy.data_ptr(), d_basic.data_ptr(), noisy.shape, fp, bp, nparts,
                                 stride, p2.beta, x.data_ptr(), t.data_ptr(), fl.data_ptr(), max_rows)
    keep = fl[:rows].bool()
    return x[:rows][keep].contiguous(), t[:rows][keep].contiguous()


def train(ctx, x, t, steps=3000, batch=4096, lr=1e-3, seed=1, theta=None, log=None):
    """Adam on the device data set (x, t); returns the per-step losses."""
    ctx.mlp_set(init_theta(seed) if theta is None else theta)
    losses = []
    done = 0
    while done < steps:
        n = min(500, steps - done)
        losses.extend(ctx.mlp_train_dev(x.data_ptr(), t.data_ptr(), x.shape[0], min(batch, x.shape[0]), n,
                                        lr, seed))
        done += n
        if log:
            log(f"step {done}: batch loss {np.mean(losses[-n:]):.5f}")
    return np.array(losses)


def evaluate(ctx, A, S, theta, n=20, size=(960, 540, 20), sigmas=(10.0, 20.0, 30.0, 40.0), seed0=9000,
             nparts=2, threads=1, log=None):
    """Held-out evaluation of a network: n seeded sequences (seeds from
    seed0, outside the training range) of size W x H x F at the given noise
    levels in turn, each with TV-L1 flows (5 scales, 10 warps, lambda 0.15)
    and the plain parameters px = pt = 4, np 32, 11 x 11 search.  Per
    sequence: PSNR of the plain two-step output and of the learned single
    pass, and the device time of each (inputs resident).  Returns
    (per-sequence list, summary)."""
    import torch
    W, H, F = size
    dev = torch.device("cuda", ctx.device)
    ctx.mlp_set(theta)
    psnr = lambda v, c: float(10 * np.log10(255.0 ** 2 / np.mean((v.astype(np.float64) - c) ** 2)))
    res = []
    for i in range(n):
        sigma = float(sigmas[i % len(sigmas)])
        clean = S.video(W, H, F, seed=seed0 + i, threads=threads)
        noisy = S.add_noise(clean, sigma, seed0 + 50000 + i)
        dn = torch.from_numpy(noisy).to(dev)
        ff = torch.empty((F, 2, H, W), dtype=torch.float32, device=dev)
        bf = torch.empty_like(ff)
        ctx.tvl1_sequence_dev(dn.data_ptr(), W, H, F, ff.data_ptr(), bf.data_ptr())
        p1 = A.cli_params(1, sigma, noisy.shape, px=4, pt=4, wx=11, np_=32)
        p2 = A.cli_params(2, sigma, noisy.shape, px=4, pt=4, wx=11, np_=32)
        db, dfin, dl = torch.empty_like(dn), torch.empty_like(dn), torch.empty_like(dn)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.run_dev(p1, p2, dn.data_ptr(), noisy.shape, ff.data_ptr(), bf.data_ptr(), db.data_ptr(),
                    dfin.data_ptr(), nparts=nparts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ctx.run_learned_dev(p1, p2, dn.data_ptr(), noisy.shape, ff.data_ptr(), bf.data_ptr(), dl.data_ptr(),
                            nparts=nparts)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        r = {"seed": seed0 + i, "sigma": sigma, "psnr_noisy": psnr(noisy, clean),
             "psnr_plain": psnr(dfin.cpu().numpy(), clean), "psnr_learned": psnr(dl.cpu().numpy(), clean),
             "ms_plain": (t1 - t0) * 1e3, "ms_learned": (t2 - t1) * 1e3}
        res.append(r)
        if log:
            log(f"held-out {i + 1}/{n} sigma {sigma:g}: plain {r['psnr_plain']:.2f} dB, learned "
                f"{r['psnr_learned']:.2f} dB, {r['ms_plain']:.1f} / {r['ms_learned']:.1f} ms")
    # the first sequence also plans and allocates: speeds from the others
    timed = res[1:] if len(res) > 1 else res
    summary = {
        "sequences": n, "size": f"{W}x{H}x{F}", "sigmas": list(sigmas),
        "dpsnr_mean_db": float(np.mean([r["psnr_learned"] - r["psnr_plain"] for r in res])),
        "dpsnr_by_sigma_db": {str(sg): float(np.mean([r["psnr_learned"] - r["psnr_plain"] for r in res
                                                      if r["sigma"] == sg])) for sg in sigmas},
        "speedup_mean": float(np.mean([r["ms_plain"] / r["ms_learned"] for r in timed])),
    }
    return res, summary


def main():
    import importlib.util
    import json
    import sys
    import torch
    ap = argparse.ArgumentParser(description="train the gain network of the learned single pass")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--sequences", type=int, default=24)
    ap.add_argument("--size", default="480x272x12", help="WxHxF of each training sequence")
    ap.add_argument("--sigmas", default="10,20,30,40")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--stride", type=int, default=4)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--eval", type=int, default=0, metavar="N",
                    help="after training (or with --steps 0, for the weights in --out), evaluate on N "
                         "held-out 960x540x20 sequences against the plain two-step path")
    ap.add_argument("--threads", type=int, default=1, help="threads of the scene generator")
    args = ap.parse_args()
    pkg = sys.modules.get(__package__) if __package__ else None
    if pkg is None:
        spec = importlib.util.spec_from_file_location("bm4dof_amd", os.path.join(HERE, "__init__.py"),
                                                      submodule_search_locations=[HERE])
        pkg = importlib.util.module_from_spec(spec)
        sys.modules["bm4dof_amd"] = pkg
        spec.loader.exec_module(pkg)
    spec = importlib.util.spec_from_file_location("bm4dof_amd.synth", os.path.join(HERE, "synth.py"))
    S = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(S)
    W, H, F = (int(v) for v in args.size.lower().split("x"))
    sigmas = [float(s) for s in args.sigmas.split(",")]
    ctx = pkg.Context(0)
    if args.steps == 0:
        res, summary = evaluate(ctx, pkg, S, load(args.out), n=args.eval, threads=args.threads,
                                log=lambda m: print("[eval] " + m, flush=True))
        print(json.dumps(summary), flush=True)
        return
    xs, ts = [], []
    t0 = time.time()
    for i in range(args.sequences):
        seed = args.seed0 + i
        sigma = sigmas[i % len(sigmas)]
        clean = S.video(W, H, F, seed=seed, threads=args.threads)
        noisy = S.add_noise(clean, sigma, seed + 50000)
        x, t = training_pairs(ctx, pkg, noisy, sigma, stride=args.stride)
        xs.append(x)
        ts.append(t)
        print(f"[train] sequence {i + 1}/{args.sequences} sigma {sigma:g}: {x.shape[0]} rows "
              f"({time.time() - t0:.1f} s)", flush=True)
    x = torch.cat(xs)
    t = torch.cat(ts)
    del xs, ts
    losses = train(ctx, x, t, steps=args.steps, batch=args.batch, lr=args.lr,
                   log=lambda m: print("[train] " + m, flush=True))
    theta = ctx.mlp_get()
    save(theta, args.out)
    print(f"[train] {x.shape[0]} rows, final loss {losses[-100:].mean():.5f}, saved {args.out}", flush=True)
    if args.eval:
        del x, t
        res, summary = evaluate(ctx, pkg, S, theta, n=args.eval, threads=args.threads,
                                log=lambda m: print("[eval] " + m, flush=True))
        print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
