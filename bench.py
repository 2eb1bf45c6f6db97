#!/usr/bin/env python3
"""Throughput of the plain two-step BM4D-OF denoiser on one MI355X.

Workload: grayscale video 960x540x20 float32, sigma 20, seeded procedural
scene (20 textured shapes under random affine motion over a smooth
gradient), px=4, pt=4, np1=np2=32, 11x11 spatial search, temporal +-4 frames
along TV-L1 flow (5 scales, 10 warps, lambda 0.15).  The command line it
replaces is `bm4d-of -px1 4 -pt1 4 -px2 4 -pt2 4 -np1 32 -np2 32 -wx1 11
-wx2 11` with flows from `tvl1flow ... 0.25 0.15 0.3 5 0 0.5 10 0.01`.

One step = step 1 + step 2 (runNlBayes, src/core/nldct.cpp:267-478) over the
whole video, inputs and flows already resident in HBM (the flows are
computed once beforehand by the GPU TV-L1; its time is reported apart).
`value` is Mpix/s = 960*540*20 / seconds per step.  The `learned` entry
times the learned single pass (shipped network weights) on the same video
and flows and gives its PSNR against the plain two-step output.  The
`color_fallback` entry runs the colour workload (RGB 1920x1080x30, sigma 40,
np 48, 15x15 search, +-6 frames): plain two-step, learned single pass, and
learned pass with the plain fallback, with the fallback fraction.

The roofline entry is for the stage with the largest share of the step
(its main kernel named), from the library's own HIP events on its stream;
`roofline_second` is the next stage.  The CPU baseline times the
C restatement of the reference (oracle/) on a bounded crop of the same
video with OpenMP over tiles (nParts = 2 x threads, as the reference).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
W, H, F, SIGMA = 960, 540, 20, 20.0
METRIC = "denoised Mpix/s at 960x540x20 σ=20, plain vs learned; ΔPSNR vs CPU ref"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def _load(name, rel):
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(REPO, rel)
    kw = {}
    if rel.endswith("__init__.py"):
        kw["submodule_search_locations"] = [os.path.dirname(path)]
    spec = importlib.util.spec_from_file_location(name, path, **kw)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def _psnr(a, b):
    import numpy as np
    d = a.astype(np.float64) - b.astype(np.float64)
    return float(10 * np.log10(255.0 ** 2 / np.mean(d * d)))


def _stage_work(stats, anchors, pixels, P=64, C=1):
    """Work one launch of each stage's main kernel executes, as (bound,
    amount, basis).  coef_kernel: the forward transform of every patch
    anchor, 2 A P^2 FLOP (A = anchors).  Group filters (fp32 VALU, no matrix
    cores): the forward transforms come from those tables, so a filter
    launch executes the inverse transform, 2 N P^2, the Haar transforms and
    the shrinkage: step 1 + 8 P (N - 1) + 2 P N, step 2 + 12 P (N - 1) + 5 P N
    (N = group rows, the counts of the reference's definitions).  Ordered
    aggregation: bytes, each group row's estimate and weight read and the
    numerator and weight sums written once, N (4 P C + 4) + 4 pixels (C + 1)."""
    n1, n2, g1, g2 = stats["rows1"], stats["rows2"], stats["groups1"], stats["groups2"]
    coef = ("valu", 2.0 * anchors * P * P * C, "forward transform of every patch anchor, 2 A P^2")
    f1 = 2.0 * n1 * P * P * C + 8.0 * P * C * (n1 - g1) + 2.0 * P * C * n1
    f2 = 2.0 * n2 * P * P * C + 12.0 * P * C * (n2 - g2) + 5.0 * P * C * n2
    # block matching: the sequential distance of every candidate, a
    # subtraction, a multiply and an add per patch sample (the windows of
    # 11 x 11 x 9 candidates; border windows are smaller, so an upper bound)
    cand = 11 * 11 * 9
    bm1 = ("valu", 3.0 * P * cand * g1, "distances 3 P Cand per group, Cand = 11 x 11 x 9")
    bm2 = ("valu", 3.0 * P * cand * g2, "distances 3 P Cand per group, Cand = 11 x 11 x 9")
    return {"s1_coef": coef, "s2_coef": coef, "s1_bm": bm1, "s2_bm": bm2,
            "s1_filter": ("valu", f1, "executed: inverse transform 2 N P^2 + Haar 8 P (N - 1) + threshold 2 P N"),
            "s2_filter": ("valu", f2, "executed: inverse transform 2 N P^2 + Haar 12 P (N - 1) + Wiener 5 P N"),
            "s1_aggregate": ("hbm", n1 * (4.0 * P * C + 4) + 4.0 * pixels * (C + 1),
                             "estimate rows + weights read, sums written once"),
            "s2_aggregate": ("hbm", n2 * (4.0 * P * C + 4) + 4.0 * pixels * (C + 1),
                             "estimate rows + weights read, sums written once")}


def _step_flop(stats, cand, P=64):
    """The denoiser's algorithmic work per video by the reference's
    definitions (block matching 3 P Cand per group, dense transforms
    4 N P^2 / 6 N P^2, Haar, shrinkage, aggregation 5 N P)."""
    n1, n2, g1, g2 = stats["rows1"], stats["rows2"], stats["groups1"], stats["groups2"]
    bm = 3.0 * P * cand * (g1 + g2)
    tr = 4.0 * n1 * P * P + 6.0 * n2 * P * P
    rest = 8.0 * P * (n1 - g1) + 2.0 * P * n1 + 12.0 * P * (n2 - g2) + 5.0 * P * n2 + 5.0 * P * (n1 + n2)
    return bm + tr + rest


def _roofline(stage, stage_ms, launches, steps, work, kern):
    avg_ms = stage_ms[stage] / launches[stage]
    if stage not in work:
        return {"bound": "hbm", "achieved": None, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": None,
                "traffic": _traffic(kern[stage]), "kernel": kern[stage], "stage": stage,
                "avg_launch_ms": round(avg_ms, 3)}
    bound, amount, basis = work[stage]
    per_launch = amount / (launches[stage] / steps)
    if bound == "valu":
        ach = per_launch / (avg_ms * 1e-3) / 1e12
        peak, unit, key = PEAK_FP32_TFLOPS, "TFLOP/s", "flop_per_launch"
    else:
        ach = per_launch / (avg_ms * 1e-3) / 1e9
        peak, unit, key = PEAK_HBM_GBS, "GB/s", "bytes_per_launch"
    return {"bound": bound, "achieved": round(ach, 3), "peak": peak, "unit": unit,
            "frac": round(ach / peak, 4), "traffic": _traffic(kern[stage]), "kernel": kern[stage],
            "stage": stage, "avg_launch_ms": round(avg_ms, 3), key: per_launch, "basis": basis}


def _traffic(kernel):
    """HBM bytes per launch of `kernel`, as committed from rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE runs scaled by a copy-kernel calibration
    (profiles/traffic.txt: tab-separated lines `kernel fetch_bytes write_bytes`)."""
    path = os.path.join(REPO, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("#") or "\t" not in line:
            continue
        f = line.rstrip("\n").split("\t")
        if len(f) == 3 and f[0] == kernel:
            return float(f[1]) + float(f[2])
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(noisy, clean, ff, bf, crop, threads, gpu_run, reps=3, tv_pairs=0):
    """Time the oracle (C restatement of nldct.cpp, scalar per thread) on a
    crop of the workload, the median of `reps` runs; returns the bench's
    cpu_baseline object and the comparison of the GPU output on the same crop
    and tiling (gpu_run) with the oracle's: max-abs difference and PSNR
    difference of the final estimates.  With tv_pairs > 0 the oracle's TV-L1
    (single-threaded, as the reference's with one thread) is also timed on
    that many full-size pairs, one per thread, and the 38 solves of the
    sequence extrapolated from it."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, REPO)
    import oracle as O
    cw, ch = crop
    x0, y0 = (W - cw) // 2, (H - ch) // 2
    sl = (slice(None), slice(None), slice(y0, y0 + ch), slice(x0, x0 + cw))
    n = np.ascontiguousarray(noisy[sl])
    f1 = np.ascontiguousarray(ff[sl])
    b1 = np.ascontiguousarray(bf[sl])
    o1 = O.cli_params(1, SIGMA, F, px=4, pt=4, wx=11, np_=32)
    o2 = O.cli_params(2, SIGMA, F, px=4, pt=4, wx=11, np_=32)
    O._omp_threads(threads)
    times = []
    for r in range(reps):
        t0 = time.perf_counter()
        _, cpu_final, _ = O.run(o1, o2, n, f1, b1, nparts=2 * threads)
        times.append(time.perf_counter() - t0)
        print(f"[bench] CPU oracle run {r + 1}/{reps}: {times[-1]:.1f} s", file=sys.stderr, flush=True)
    dt = float(np.median(times))
    gpu_final = gpu_run(n, f1, b1, 2 * threads)
    c = clean[sl]
    parity = {"sample": f"{cw}x{ch}x{F} crop, nParts={2 * threads}",
              "max_abs_diff": float(np.abs(gpu_final.astype(np.float64) - cpu_final).max()),
              "identical": bool(np.array_equal(gpu_final, cpu_final)),
              "dpsnr_db": round(_psnr(gpu_final, c) - _psnr(cpu_final, c), 6),
              "psnr_cpu_db": round(_psnr(cpu_final, c), 3)}
    out = {"value": round(cw * ch * F / dt / 1e6, 5), "unit": "Mpix/s", "cores": threads,
           "kind": "port",
           "sample": f"{cw}x{ch}x{F} centre crop of the same video and flows, both steps, "
                     f"{threads} OpenMP threads of {os.cpu_count()} ({_cpu_model()}), "
                     f"nParts={2 * threads}, median of {reps}: {dt:.1f} s "
                     f"(runs {', '.join(f'{t:.1f}' for t in times)} s); the C restatement of the "
                     "reference with a sequential-K fused-multiply-add sgemm (the reference itself needs "
                     "cblas.h and fftw3.h, absent here)"}
    if tv_pairs > 0:
        pairs = [(i, i + 1) for i in range(tv_pairs)]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(min(threads, tv_pairs)) as ex:
            list(ex.map(lambda ij: O.tvl1(noisy[ij[0], 0], noisy[ij[1], 0], lam=0.15), pairs))
        tv_s = time.perf_counter() - t0
        per_pair = tv_s * min(threads, tv_pairs) / tv_pairs
        seq_s = per_pair * 2 * (F - 1) / threads
        out["tvl1"] = {"pairs_timed": tv_pairs, "wall_s": round(tv_s, 2),
                       "s_per_pair_one_thread": round(per_pair, 2),
                       "sequence_38_pairs_s": round(seq_s, 2),
                       "sample": f"{tv_pairs} full-size 960x540 pairs, one per thread "
                                 f"(5 scales, 10 warps, lambda 0.15); the 38 solves of the sequence "
                                 f"at {threads} pairs at a time"}
        full_denoise_s = dt * (W * H) / (cw * ch)
        out["e2e_value"] = round(W * H * F / (full_denoise_s + seq_s) / 1e6, 5)
    return out, parity


def batch32(A, S, ctx, torch, nparts):
    """32 gray sequences, seeded: frames uniform in [3, 30], size from
    {320x240, 640x360, 960x540, 1280x720}, sigma from {10, 20, 40}, px=pt=4,
    np 32, 11x11 search; both steps of all of them in one captured hipGraph,
    options picked per shape class by the autotuner.  Flow: none (the batch
    measures the denoiser)."""
    import numpy as np
    rng = np.random.default_rng(2024)
    sizes = [(320, 240), (640, 360), (960, 540), (1280, 720)]
    seqs, keep, p1s, p2s = [], [], [], []
    pix = 0
    t0 = time.perf_counter()
    for i in range(32):
        w, h = sizes[rng.integers(0, 4)]
        f = int(rng.integers(3, 31))
        sigma = float(rng.choice([10.0, 20.0, 40.0]))
        clean = S.video(w, h, f, seed=7000 + i)
        noisy = S.add_noise(clean, sigma, 8000 + i)
        dn = torch.from_numpy(noisy).cuda()
        db, df = torch.empty_like(dn), torch.empty_like(dn)
        keep += [dn, db, df]
        seqs.append(A.Seq(dn.data_ptr(), None, None, db.data_ptr(), df.data_ptr(), w, h, f))
        p1s.append(A.cli_params(1, sigma, noisy.shape, px=4, pt=4, wx=11, np_=32))
        p2s.append(A.cli_params(2, sigma, noisy.shape, px=4, pt=4, wx=11, np_=32))
        pix += w * h * f
        print(f"[bench] batch sequence {i + 1}/32: {w}x{h}x{f} sigma {sigma:g}", file=sys.stderr, flush=True)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch = ctx.batch(seqs, p1s, p2s, nparts=nparts, graph=True, autotune=True)
    create_s = time.perf_counter() - t0
    batch.run()   # warm
    runs = [batch.run() for _ in range(3)]
    batch.close()
    ms = float(np.median(runs))
    # the same sequences one at a time through the single-video entry point
    # (default options, each run once to warm the buffers, then timed)
    single_ms = 0.0
    for q, p1, p2 in zip(seqs, p1s, p2s):
        shape = (q.f, 1, q.h, q.w)

        def one():
            ctx.run_dev(p1, p2, q.noisy, shape, 0, 0, q.basic, q.final, nparts=nparts)
        one()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        one()
        torch.cuda.synchronize()
        single_ms += (time.perf_counter() - t0) * 1e3
    return {"value": round(pix / (ms * 1e-3) / 1e6, 4), "unit": "Mpix/s", "sequences": 32,
            "mpix": round(pix / 1e6, 3), "ms_per_batch": round(ms, 3), "replays_ms": [round(r, 3) for r in runs],
            "singles_sum_ms": round(single_ms, 3), "speedup_vs_singles": round(single_ms / ms, 3),
            "create_s_incl_autotune": round(create_s, 2), "generation_s": round(gen_s, 1)}


def color_fallback(A, S, L, ctx, torch, nparts, gvar_thr, threads):
    """Colour RGB video 1920x1080x30, sigma 40, filtered in the opponent
    colour space (the parameters' default YUV transform), seeded scene of 40
    moving shapes with per-channel textures; px=pt=4, np1=np2=48 (the
    power-of-two cut keeps 32), 15x15 spatial search, temporal +-6 frames
    along TV-L1 flow computed on the luma of the noisy video.  Times the
    plain two-step path, the learned single pass, and the learned pass with
    the plain fallback under groups whose gain variance exceeds gvar_thr,
    and reports the flagged fraction and PSNRs."""
    import numpy as np
    Wc, Hc, Fc, sig = 1920, 1080, 30, 40.0
    t0 = time.perf_counter()
    clean = S.video(Wc, Hc, Fc, channels=3, n_shapes=40, seed=4242, threads=threads)
    noisy = S.add_noise(clean, sig, 4243)
    gen_s = time.perf_counter() - t0
    print(f"[bench] colour video {Wc}x{Hc}x{Fc}x3 generated in {gen_s:.1f} s", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    dn = torch.from_numpy(noisy).to(dev)
    luma = (0.299 * dn[:, 0] + 0.587 * dn[:, 1] + 0.114 * dn[:, 2]).contiguous()
    ff = torch.empty((Fc, 2, Hc, Wc), dtype=torch.float32, device=dev)
    bf = torch.empty_like(ff)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.tvl1_sequence_dev(luma.data_ptr(), Wc, Hc, Fc, ff.data_ptr(), bf.data_ptr())
    torch.cuda.synchronize()
    tvl1_s = time.perf_counter() - t0
    p1 = A.cli_params(1, sig, noisy.shape, px=4, pt=4, wx=15, np_=48)
    p2 = A.cli_params(2, sig, noisy.shape, px=4, pt=4, wx=15, np_=48)
    for p in (p1, p2):
        p.sizeSearchTimeRangeFwd = p.sizeSearchTimeRangeBwd = 6
    db, dfin, dl = torch.empty_like(dn), torch.empty_like(dn), torch.empty_like(dn)

    def timed(fn, reps):
        out = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
        return dt, out   # the last repetition (the first plans and allocates)

    plain_s, _ = timed(lambda: ctx.run_dev(p1, p2, dn.data_ptr(), noisy.shape, ff.data_ptr(), bf.data_ptr(),
                                           db.data_ptr(), dfin.data_ptr(), nparts=nparts), 2)
    plain_st = {}
    for name, ms in ctx.timings():
        plain_st[name] = round(plain_st.get(name, 0.0) + ms, 3)
    plain = dfin.cpu().numpy()
    print(f"[bench] colour plain two-step: {plain_s:.3f} s", file=sys.stderr, flush=True)
    ctx.mlp_set(L.load())

    def learned(thr):
        return ctx.run_learned_dev(p1, p2, dn.data_ptr(), noisy.shape, ff.data_ptr(), bf.data_ptr(),
                                   dl.data_ptr(), nparts=nparts, gvar_thr=thr)

    l_s, _ = timed(lambda: learned(-1.0), 2)
    l_psnr = _psnr(dl.cpu().numpy(), clean)
    fb_s, frac = timed(lambda: learned(gvar_thr), 1)
    fb_psnr = _psnr(dl.cpu().numpy(), clean)
    print(f"[bench] colour learned {l_s:.3f} s, with fallback {fb_s:.3f} s (flagged {frac:.4f})",
          file=sys.stderr, flush=True)
    mpix = Wc * Hc * Fc / 1e6
    return {
        "workload": f"colour RGB {Wc}x{Hc}x{Fc}, sigma 40, opponent colour space, px 4, pt 4, np1=np2=48, "
                    "wx 15, temporal +-6 along TV-L1 flow (luma)",
        "unit": "Mpix/s",
        "plain": {"value": round(mpix / plain_s, 4), "s": round(plain_s, 3),
                  "psnr_db": round(_psnr(plain, clean), 3), "stages_ms": plain_st},
        "learned": {"value": round(mpix / l_s, 4), "s": round(l_s, 3), "psnr_db": round(l_psnr, 3)},
        "learned_with_fallback": {"value": round(mpix / fb_s, 4), "s": round(fb_s, 3),
                                  "psnr_db": round(fb_psnr, 3), "gvar_thr": gvar_thr,
                                  "fallback_fraction": round(frac, 5)},
        "psnr_noisy_db": round(_psnr(noisy, clean), 3),
        "tvl1_sequence_s": round(tvl1_s, 3), "generation_s": round(gen_s, 1),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nparts", type=int, default=2,
                    help="spatial tiles (the reference uses 2 x OpenMP threads)")
    ap.add_argument("--cpu-crop", default="480x270", help="crop WxH of the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, available cores)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned single pass")
    ap.add_argument("--no-batch32", action="store_true",
                    help="skip the batched workload (32 uneven gray sequences in one hipGraph)")
    ap.add_argument("--cpu-reps", type=int, default=3, help="CPU baseline runs (median)")
    ap.add_argument("--cpu-sweep", default="",
                    help="comma-separated thread counts (e.g. 1,8,16): time the CPU baseline at each "
                         "(nParts = 2 T, median of --cpu-reps) and report the fastest as cpu_baseline")
    ap.add_argument("--cpu-tvl1-pairs", type=int, default=16,
                    help="full-size pairs of the CPU TV-L1 sample (0: none)")
    ap.add_argument("--learned-eval", type=int, default=0, metavar="N",
                    help="also evaluate the shipped network on N held-out 960x540x20 sequences")
    ap.add_argument("--no-color", action="store_true",
                    help="skip the colour 1920x1080x30 learned pass with plain fallback")
    ap.add_argument("--color-gvar", type=float, default=0.02,
                    help="gain-variance threshold of the colour fallback")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the denoiser "
              "has no multi-GPU path in this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if args.copy_calibration:
        # a copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch scale
        # the counters of the denoiser's kernels to bytes
        n = 1 << 28
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}), flush=True)
        return
    A = _load("bm4dof_amd", os.path.join("bm4d-of_amd", "__init__.py"))
    S = _load("bm4dof_amd.synth", os.path.join("bm4d-of_amd", "synth.py"))
    A.lib()
    print(f"[bench] workload {W}x{H}x{F} sigma {SIGMA}", file=sys.stderr, flush=True)

    clean = S.video(W, H, F, seed=101)
    noisy = S.add_noise(clean, SIGMA, 201)
    dev = torch.device("cuda:0")
    d_noisy = torch.from_numpy(noisy).to(dev)
    d_ff = torch.empty((F, 2, H, W), dtype=torch.float32, device=dev)
    d_bf = torch.empty_like(d_ff)
    d_basic = torch.empty_like(d_noisy)
    d_final = torch.empty_like(d_noisy)
    torch.cuda.synchronize()
    ctx = A.Context(0)
    # a caller that denoises videos of one shape repeatedly keeps the
    # coefficient tables between calls (the default frees them at each end)
    ctx.set_option("keep_tables", 1)

    # flows of the sequence (a first call sizes the buffers; the second is timed)
    ctx.tvl1_sequence_dev(d_noisy.data_ptr(), W, H, F, d_ff.data_ptr(), d_bf.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.tvl1_sequence_dev(d_noisy.data_ptr(), W, H, F, d_ff.data_ptr(), d_bf.data_ptr())
    torch.cuda.synchronize()
    tvl1_s = time.perf_counter() - t0
    tvl1_pix_it = ctx.stats().get("tvl1_pixel_iters", 0)
    print(f"[bench] TV-L1 flows of {2 * (F - 1)} pairs: {tvl1_s:.3f} s", file=sys.stderr, flush=True)

    p1 = A.cli_params(1, SIGMA, noisy.shape, px=4, pt=4, wx=11, np_=32)
    p2 = A.cli_params(2, SIGMA, noisy.shape, px=4, pt=4, wx=11, np_=32)

    def step():
        ctx.run_dev(p1, p2, d_noisy.data_ptr(), noisy.shape, d_ff.data_ptr(), d_bf.data_ptr(),
                    d_basic.data_ptr(), d_final.data_ptr(), nparts=args.nparts)

    for i in range(args.warmup):
        step()
        print(f"[bench] warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()

    stage_ms, launches = {}, {}
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        for name, ms in ctx.timings():
            stage_ms[name] = stage_ms.get(name, 0.0) + ms
            launches[name] = launches.get(name, 0) + 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stats = ctx.stats()
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    ms_per_step = elapsed / args.steps * 1e3
    value = W * H * F / (elapsed / args.steps) / 1e6
    basic = d_basic.cpu().numpy()
    final = d_final.cpu().numpy()

    # dominant kernel among the transform/filter, matching and aggregation stages
    # kernels this workload runs (one channel, 4x4x4 patches, 32 neighbours)
    kern = {"s1_bm": "bm_reg_kernel<4, 4, false>", "s2_bm": "bm_reg_kernel<4, 4, false>",
            "s1_coef": "coef_kernel<4>", "s2_coef": "coef_kernel<4>",
            "s1_filter": "filter1_fast_kernel<4>", "s2_filter": "filter2_fast_kernel<4>",
            "s1_aggregate": "agg_bin_fold_kernel<1, 32, 16, 4>", "s2_aggregate": "agg_bin_fold_kernel<1, 32, 16, 4>"}
    work = _stage_work(stats, (W - 3) * (H - 3) * (F - 3), W * H * F)
    ranked = sorted((k for k in stage_ms if k in kern), key=lambda k: -stage_ms[k])
    roofline = _roofline(ranked[0], stage_ms, launches, args.steps, work, kern)
    roofline_next = _roofline(ranked[1], stage_ms, launches, args.steps, work, kern)
    step_flop = _step_flop(stats, 11 * 11 * 9)
    step_tf = step_flop / (ms_per_step * 1e-3) / 1e12
    step_roofline = {"bound": "valu", "flop_per_step": step_flop, "achieved": round(step_tf, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(step_tf / PEAK_FP32_TFLOPS, 4),
                     "basis": "the reference's algorithmic work per video: block matching 3 P Cand per group, "
                              "dense transforms 4 N P^2 and 6 N P^2, Haar, shrinkage, aggregation 5 N P"}

    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural video, i.i.d. Gaussian noise)",
        "config": {"workload": f"plain two-step BM4D-OF, gray video {W}x{H}x{F}, sigma 20, "
                               "px 4, pt 4, np1=np2=32, wx 11, temporal +-4 along TV-L1 flow",
                   "nparts": args.nparts, "path": "plain"},
        "roofline": roofline,
        "roofline_second": roofline_next,
        "step_roofline": step_roofline,
        "stages_ms_per_step": {k: round(v / args.steps, 3) for k, v in stage_ms.items()},
        "tvl1_sequence_s": round(tvl1_s, 3),
        # TV-L1 iterations (tvl1flow_lib.c:165-246) over the executed
        # pixel-iterations: about 60 FLOP each (TH step, divergences, u, the
        # squared update, forward gradients, two hypot, four divisions; a
        # division or square root counted as one), the whole flow computation
        # timed; and the bytes an iteration that keeps nothing on chip moves,
        # 64 B (u1, u2, p11..p22 read and written, rho_c, I1wx, I1wy, grad
        # read), which the LDS tiles of 4 iterations cut to about a third
        "tvl1_roofline": {"bound": "valu", "pixel_iterations": int(tvl1_pix_it), "flop_per_pixel_iteration": 60,
                          "achieved": round(60.0 * tvl1_pix_it / tvl1_s / 1e12, 3), "peak": PEAK_FP32_TFLOPS,
                          "unit": "TFLOP/s", "frac": round(60.0 * tvl1_pix_it / tvl1_s / 1e12 / PEAK_FP32_TFLOPS, 4),
                          "untiled_equivalent_gbs": round(64.0 * tvl1_pix_it / tvl1_s / 1e9, 1)},
        "e2e_mpix_s": round(W * H * F / (elapsed / args.steps + tvl1_s) / 1e6, 4),
        "psnr_db": {"noisy": round(_psnr(noisy, clean), 3), "basic": round(_psnr(basic, clean), 3),
                    "final": round(_psnr(final, clean), 3)},
        "groups": {"step1": stats["groups1"], "step2": stats["groups2"]},
    }
    # learned single pass (shipped weights): the same video and flows
    if not args.no_learned:
        L = _load("bm4dof_amd.learned", os.path.join("bm4d-of_amd", "learned.py"))
        ctx.mlp_set(L.load())
        d_learned = torch.empty_like(d_noisy)

        def lstep():
            return ctx.run_learned_dev(p1, p2, d_noisy.data_ptr(), noisy.shape, d_ff.data_ptr(),
                                       d_bf.data_ptr(), d_learned.data_ptr(), nparts=args.nparts)

        for _ in range(args.warmup):
            lstep()
        torch.cuda.synchronize()
        lstage = {}
        t0 = time.perf_counter()
        for _ in range(args.steps):
            lstep()
            for name, ms in ctx.timings():
                lstage[name] = lstage.get(name, 0.0) + ms
        torch.cuda.synchronize()
        lel = time.perf_counter() - t0
        learned = d_learned.cpu().numpy()
        lv = W * H * F / (lel / args.steps) / 1e6
        out["learned"] = {
            "value": round(lv, 4), "unit": "Mpix/s", "ms_per_step": round(lel / args.steps * 1e3, 3),
            "speedup_vs_plain": round(lv / value, 3),
            "psnr_db": round(_psnr(learned, clean), 3),
            "dpsnr_vs_plain_db": round(_psnr(learned, clean) - out["psnr_db"]["final"], 3),
            "stages_ms_per_step": {k: round(v / args.steps, 3) for k, v in lstage.items()},
            "weights": "bm4d-of_amd/weights/gain_mlp_66_128_128_64.f32",
        }
        print(f"[bench] learned single pass: {lv:.2f} Mpix/s", file=sys.stderr, flush=True)
    if not args.no_batch32:
        out["batch32"] = batch32(A, S, ctx, torch, args.nparts)
    if args.learned_eval:
        L = _load("bm4dof_amd.learned", os.path.join("bm4d-of_amd", "learned.py"))
        _, out["learned_heldout"] = L.evaluate(ctx, A, S, L.load(), n=args.learned_eval,
                                               threads=min(16, len(os.sched_getaffinity(0))))
    if not args.no_color:
        L = _load("bm4dof_amd.learned", os.path.join("bm4d-of_amd", "learned.py"))
        out["color_fallback"] = color_fallback(A, S, L, ctx, torch, args.nparts, args.color_gvar,
                                               min(16, len(os.sched_getaffinity(0))))
    if not args.no_cpu:
        cw, ch = (int(v) for v in args.cpu_crop.lower().split("x"))
        thr = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        print(f"[bench] CPU baseline: oracle on {cw}x{ch}x{F}, {thr} threads", file=sys.stderr,
              flush=True)

        def gpu_run(n, f1, b1, nparts):
            q1 = A.cli_params(1, SIGMA, n.shape, px=4, pt=4, wx=11, np_=32)
            q2 = A.cli_params(2, SIGMA, n.shape, px=4, pt=4, wx=11, np_=32)
            return ctx.run(q1, q2, n, f1, b1, nparts=nparts)[1]

        h_ff, h_bf = d_ff.cpu().numpy(), d_bf.cpu().numpy()
        if args.cpu_sweep:
            sweep = []
            for t in (int(v) for v in args.cpu_sweep.split(",")):
                print(f"[bench] CPU sweep: {t} threads", file=sys.stderr, flush=True)
                cb, par = cpu_baseline(noisy, clean, h_ff, h_bf, (cw, ch), t, gpu_run, reps=args.cpu_reps,
                                       tv_pairs=args.cpu_tvl1_pairs if t == thr else 0)
                sweep.append((cb, par))
            out["cpu_sweep"] = [{"threads": cb["cores"], "nparts": 2 * cb["cores"], "value": cb["value"],
                                 "sample": cb["sample"], "identical_to_gpu": par["identical"]}
                                for cb, par in sweep]
            out["cpu_baseline"], out["parity_vs_cpu"] = max(sweep, key=lambda q: q[0]["value"])
        else:
            out["cpu_baseline"], out["parity_vs_cpu"] = cpu_baseline(
                noisy, clean, h_ff, h_bf, (cw, ch), thr, gpu_run,
                reps=args.cpu_reps, tv_pairs=args.cpu_tvl1_pairs)
    else:
        out["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
