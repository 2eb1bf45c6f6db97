"""Float64 restatement of the reference denoiser for tiny inputs.

TEST INFRASTRUCTURE ONLY: a second statement of pariasm/bm4d-of's
runNlBayes / processNlBayes (src/core/nldct.cpp:267-2448) written directly
from the reference text in numpy, beside the C restatement in
oracle/nldct_oracle.c.  The pin tests compare the two on small videos; a
misreading of the reference that only one of them makes (a trajectory read at
the wrong point, another rounding, an off-by-one in a grid border, a channel
left out of a distance, a weight not left stale) shows up as a difference.

What is deliberately different from the oracle, so that the two statements
do not share code paths:
  * arithmetic in float64 (transforms, filters, sums), with the patch
    transform as one matrix product and the group transform as the textbook
    orthonormal Haar matrix rather than the in-place butterflies;
  * the DCT basis from scipy, the Kaiser windows from numpy.kaiser;
  * candidate order under equal distances from libstdc++'s own
    std::partial_sort, compiled from a few lines below and called per group.

The patch distances, which decide the neighbour lists, are formed as the
reference forms them (a float32 running sum in its loop order), and so is the
colour transform whose outputs they are taken on: these discrete decisions
must come out the same.  Where float64 and the reference's float32 could
still decide a discontinuous step differently (a coefficient at the hard
threshold), the case is counted in `Mini.ambiguous`.
"""
import ctypes as C
import os
import shutil
import subprocess
import tempfile

import numpy as np
import scipy.fft

_PSORT_SRC = r"""
#include <algorithm>
#include <utility>
#include <vector>
static bool cmp(const std::pair<float, unsigned> &a, const std::pair<float, unsigned> &b)
{ return a.first < b.first; }
extern "C" void mini_partial_sort(const float *d, long n, long m, unsigned *out)
{
  std::vector<std::pair<float, unsigned> > v(n);
  for (long i = 0; i < n; i++) { v[i].first = d[i]; v[i].second = (unsigned)i; }
  std::partial_sort(v.begin(), v.begin() + m, v.end(), cmp);
  for (long i = 0; i < m; i++) out[i] = v[i].second;
}
"""
_PSORT = None


def _psort_lib():
    """libstdc++'s std::partial_sort with the reference's comparator
    (comparaisonFirst, src/utils/utilities.cpp:35-40)."""
    global _PSORT
    if _PSORT is None:
        gxx = shutil.which("g++")
        if gxx is None:
            raise RuntimeError("g++ is needed for the libstdc++ partial_sort helper")
        d = tempfile.mkdtemp(prefix="mini_psort_")
        src = os.path.join(d, "ps.cpp")
        so = os.path.join(d, "ps.so")
        with open(src, "w") as f:
            f.write(_PSORT_SRC)
        subprocess.check_call([gxx, "-O1", "-shared", "-fPIC", "-o", so, src])
        L = C.CDLL(so)
        L.mini_partial_sort.argtypes = [np.ctypeslib.ndpointer(np.float32, flags="C"), C.c_long,
                                        C.c_long, np.ctypeslib.ndpointer(np.uint32, flags="C")]
        _PSORT = L
    return _PSORT


def partial_sort_order(d32, m):
    """Positions of the m smallest of d32 in std::partial_sort's order."""
    d32 = np.ascontiguousarray(d32, np.float32)
    out = np.zeros(m, np.uint32)
    _psort_lib().mini_partial_sort(d32, d32.size, m, out)
    return out


# --------------------------------------------------------------- helpers

def determine_factor(n):
    """determineFactor, src/utils/utilities.cpp:137-158: n = a*b, a >= b."""
    if n == 1:
        return 1, 1
    b = 2
    while n % b:
        b += 1
    a = n // b
    if b > a:
        a, b = b, n // b
    return a, b


def sub_divide(W, H, bdry, nparts):
    """subDivideTight, src/utils/lib_videot.hpp:854-916: crops
    (origin_x, origin_y, ending_x, ending_y) in tile order (rows of nW)."""
    nW, nH = determine_factor(nparts)
    wT = int(np.ceil(np.float32(W) / np.float32(nW)))
    hT = int(np.ceil(np.float32(H) / np.float32(nH)))
    crops = []
    for p in range(nH):
        for q in range(nW):
            crops.append((max(0, q * wT - bdry), max(0, p * hT - bdry),
                          min(W, (q + 1) * wT + bdry), min(H, (p + 1) * hT + bdry),
                          q * wT, p * hT, min(W, (q + 1) * wT), min(H, (p + 1) * hT)))
    return crops


def dct_1d(n):
    return scipy.fft.dct(np.eye(n), norm="ortho", axis=0)  # [k, x]


def haar(n):
    """Textbook orthonormal Haar analysis matrix, DC row first."""
    if n == 1:
        return np.ones((1, 1))
    h = haar(n // 2)
    return np.vstack([np.kron(h, [1, 1]), np.kron(np.eye(n // 2), [1, -1])]) / np.sqrt(2)


def window(px, pt, on):
    """Aggregation window (nldct.cpp:698-737): BM3D's Kaiser window (beta 2,
    four decimals) for 8x8 and 10x10 patches, else ones."""
    if on and px in (8, 10):
        k = np.kaiser(px, 2.0)
        return np.tile(np.round(np.outer(k, k), 4).ravel(), pt)
    return np.ones(px * px * pt)


def opponent(v, forward):
    """transformColorSpace for 3 channels (lib_videot.hpp:728-838) in float32,
    with the reference's constants and operation order (its outputs are
    matched on, so they must be the same bits)."""
    if v.shape[1] != 3:
        return v.copy()
    f = np.float32
    x = np.asarray(v, np.float32)
    r, g, b_ = x[:, 0], x[:, 1], x[:, 2]
    a, b = f(1) / np.sqrt(f(3)), f(1) / np.sqrt(f(2))
    if forward:
        c = f(2) * a * np.sqrt(f(2))
        out = [a * (r + g + b_), b * (r - b_),
               c * (f(0.25) * r - f(0.5) * g + f(0.25) * b_)]
    else:
        c = a / b
        out = [a * r + b * g + c * f(0.5) * b_, a * r - c * b_,
               a * r - b * g + c * f(0.5) * b_]
    return np.stack(out, axis=1).astype(np.float64)


# ------------------------------------------------------------ the method

class Mini:
    """One run of runNlBayes (nldct.cpp:267-478) on (F, C, H, W) arrays.

    p1, p2: parameter objects with the nlbParams fields (the oracle's ctypes
    struct works).  bior_bases: (B, Binv) float arrays for the bior1.5 step 1
    (the only basis not built here).  Keeps per step: the neighbour lists of
    every processed group (`groups[step]`: list of (tile, ref, index list)),
    processed group counts per tile, early-out groups that kept a stale
    weight, groups skipped by the mask clear, and the ambiguity counters.
    """

    def __init__(self, p1, p2, bior_bases=None, mistake=None, predictive=False):
        self.p = {1: p1, 2: p2}
        # the reference's VBM3D_SEARCH build (nldct.h:20)
        self.predictive = predictive
        # a planted misreading, for the tests that show the comparison with
        # the oracle detects it (None in every other use)
        self.mistake = mistake
        self.bior = bior_bases
        self.ambiguous = {"threshold": 0}
        self.groups = {1: [], 2: []}
        self.counts = {1: [], 2: []}
        self.stale = {1: 0, 2: 0}
        self.skipped = {1: 0, 2: 0}

    # ---- reference grid of a tile, nldct.cpp:513-553
    @staticmethod
    def grid(p, crop, tshape, shape, mistake=None):
        """Processing mask of a tile: (F*h*w,) bool, True at grid points."""
        F, h, w = tshape
        W, Hh, _ = shape
        sPx, sPt, sW = p.sizePatch, p.sizePatchTime, p.sizeSearchWindow
        sWt = p.sizeSearchTimeRangeFwd + p.sizeSearchTimeRangeBwd + 1
        ox, oy, ex, ey = crop[:4]
        # a border is added only where the crop does not touch the image edge
        bx0, by0, bt0 = ox > 0, oy > 0, False
        bx1, by1, bt1 = ex < W, ey < Hh, False
        ori_x = sPx - 1 + sW // 2 if bx0 else 0
        ori_y = sPx - 1 + sW // 2 if by0 else 0
        ori_f = sPt - 1 + sWt // 2 if bt0 else 0
        if mistake == "border_origin":
            ori_x += bx0
        end_x = w - (sPx - 1 + sW // 2 if bx1 else sPx - 1)
        end_y = h - (sPx - 1 + sW // 2 if by1 else sPx - 1)
        end_f = F - (sPt - 1 + sWt // 2 if bt1 else sPt - 1)
        mask = np.zeros(F * h * w, bool)
        for f in range(ori_f, end_f):
            if not ((f - ori_f) % p.offSetTime == 0 or (not bt1 and f == end_f - 1)):
                continue
            for y in range(ori_y, end_y):
                if not ((y - ori_y) % p.offSet == 0 or (not by1 and y == end_y - 1)
                        or (not by0 and y == ori_y)):
                    continue
                for x in range(ori_x, end_x):
                    if (x - ori_x) % p.offSet == 0 or (not bx1 and x == end_x - 1) \
                            or (not bx0 and x == ori_x):
                        mask[(f * h + y) * w + x] = True
        return mask

    # ---- block matching: estimateSimilarPatchesStep{1,2}, nldct.cpp:1143-1348, 1607-1789
    def match(self, img, ff, bf, p, ref, channels):
        """Neighbour list of the reference patch at ref = (x, y, t) of the
        tile `img` (F, C, h, w); distances over `channels`."""
        F, Cc, h, w = img.shape
        sPx, sPt = p.sizePatch, p.sizePatchTime
        px, py, pt = ref
        if p.nSimilarPatches <= 1:
            return [(px, py, pt)]
        if self.predictive:
            return self.match_pred(img, ff, bf, p, ref, len(channels))
        wf, wb = p.sizeSearchTimeRangeFwd, p.sizeSearchTimeRangeBwd
        sW = p.sizeSearchWindow
        # non-centred temporal range (nldct.cpp:1178-1186)
        sh = min(0, pt - wb) + max(0, pt + wf - F + sPt)
        t0, t1 = max(0, pt - wb - sh), min(F - sPt, pt + wf - sh)
        order = [pt] + list(range(pt + 1, t1 + 1)) + list(range(pt - 1, t0 - 1, -1))
        centre = {pt: (px, py)}
        # distances as the reference forms them: a float32 running sum of
        # float32 squared differences, channel, frame, row, column order
        # (nldct.cpp:1262-1274, 1723-1737)
        img32 = img.astype(np.float32)
        refp = np.ascontiguousarray(
            img32[pt:pt + sPt, channels, py:py + sPx, px:px + sPx].transpose(1, 0, 2, 3)).ravel()
        cand, pats = [], []
        for qt in order:
            if qt != pt:
                # one frame step along the flow from the neighbour frame nearer
                # to pt, read at that frame's integer centre, rounded half away
                # from zero, clamped to the tile (nldct.cpp:1219-1230)
                prev = qt - 1 if qt > pt else qt + 1
                cx, cy = centre[prev]
                fl = ff if qt > pt else bf
                if self.mistake == "flow_at_reference":
                    cx, cy = px, py
                rt = qt if self.mistake == "flow_of_target_frame" else prev
                ux = fl[rt, 0, cy, cx] if fl is not None else 0.0
                uy = fl[rt, 1, cy, cx] if fl is not None else 0.0
                if self.mistake == "flow_at_reference":
                    cx, cy = centre[prev]
                rx = np.float32(np.float32(cx) + np.float32(ux))
                ry = np.float32(np.float32(cy) + np.float32(uy))
                if self.mistake == "truncate":
                    rx, ry = np.trunc(rx), np.trunc(ry)
                elif self.mistake == "round_half_even":
                    rx, ry = np.round(rx), np.round(ry)
                else:
                    rx = np.sign(rx) * np.floor(np.abs(rx) + 0.5)
                    ry = np.sign(ry) * np.floor(np.abs(ry) + 0.5)
                lim = sPx if self.mistake == "clamp_to_last_patch" else 1
                centre[qt] = (int(min(max(rx, 0), w - lim)), int(min(max(ry, 0), h - lim)))
            cx, cy = centre[qt]
            # window shifted to fit the tile (nldct.cpp:1239-1259)
            r = (sW - 1) // 2
            shx = min(0, cx - r) + max(0, cx + r - w + sPx)
            shy = min(0, cy - r) + max(0, cy + r - h + sPx)
            x0, x1 = max(0, cx - r - shx), min(w - sPx, cx + r - shx)
            y0, y1 = max(0, cy - r - shy), min(h - sPx, cy + r - shy)
            for qy in range(y0, y1 + 1):
                for qx in range(x0, x1 + 1):
                    pats.append(img32[qt:qt + sPt, channels, qy:qy + sPx, qx:qx + sPx]
                                .transpose(1, 0, 2, 3).ravel())
                    cand.append((qx, qy, qt))
        dif = refp[None, :] - np.stack(pats)          # float32
        sq = dif * dif
        d32 = np.zeros(len(cand), np.float32)
        for j in range(sq.shape[1]):
            d32 = d32 + sq[:, j]
        nsim = min(p.nSimilarPatches, len(cand))
        sel = partial_sort_order(d32, nsim)
        tau = np.float32(p.tau) * np.float32(p.tau) * sPx * sPx * sPt  # float ops
        n = 1
        while 2 * n <= nsim and d32[sel[n]] < tau:
            n *= 2
        return [cand[i] for i in sel[:n]]

    # ---- predictive search of a VBM3D_SEARCH build, nldct.cpp:904-1130, 1368-1589
    def match_pred(self, img, ff, bf, p, ref, reps):
        """Per frame, in the schedule pt, pt+1.., pt-1..: the full window in
        the reference frame, else the union (in insertion order, as an
        std::unordered_set insertion test lets through) of npf windows of
        sizeSearchWindowPred around the flow-advanced positions of the
        previous frame's npf best; distances on channel 0, the sum repeated
        `reps` times (step 2 reads imBasic(x, y, t) inside its channel loop),
        minus dsub^2 * 255 at the reference's (x, y) in any frame; the npf
        best of each frame kept (std::partial_sort), then all kept sorted
        again and cut at a power of two below tau."""
        F, Cc, h, w = img.shape
        sPx, sPt = p.sizePatch, p.sizePatchTime
        px, py, pt = ref
        npf = p.nSimilarPatchesPred
        f32 = np.float32
        dsub = f32(f32(p.dsub) * f32(p.dsub)) * f32(255)
        wf, wb = p.sizeSearchTimeRangeFwd, p.sizeSearchTimeRangeBwd
        sh = min(0, pt - wb) + max(0, pt + wf - F + sPt)
        t0, t1 = max(0, pt - wb - sh), min(F - sPt, pt + wf - sh)
        order = [pt] + list(range(pt + 1, t1 + 1)) + list(range(pt - 1, t0 - 1, -1))
        # the planes a distance runs over, in order: channel 0, once per
        # repetition (a planted misreading reads every channel instead)
        chans = list(range(Cc)) if self.mistake == "pred_all_channels" and reps > 1 else [0] * reps
        planes = [img[:, c].astype(np.float32) for c in chans]
        refp = np.concatenate([q[pt:pt + sPt, py:py + sPx, px:px + sPx].ravel() for q in planes])
        traj = {}
        kept_d, kept_c = [], []
        for qt in order:
            if qt == pt:
                centres, sw = [(px, py)], p.sizeSearchWindow
            else:
                prev = qt - 1 if qt > pt else qt + 1
                fl = ff if qt > pt else bf
                centres, sw = [], p.sizeSearchWindowPred
                for (cx, cy) in traj[prev]:
                    ux = fl[prev, 0, cy, cx] if fl is not None else 0.0
                    uy = fl[prev, 1, cy, cx] if fl is not None else 0.0
                    rx = np.float32(np.float32(cx) + np.float32(ux))
                    ry = np.float32(np.float32(cy) + np.float32(uy))
                    rx = np.sign(rx) * np.floor(np.abs(rx) + 0.5)
                    ry = np.sign(ry) * np.floor(np.abs(ry) + 0.5)
                    centres.append((int(min(max(rx, 0), w - 1)), int(min(max(ry, 0), h - 1))))
            seen, cand, pats = set(), [], []
            r = (sw - 1) // 2
            for (cx, cy) in centres:
                shx = min(0, cx - r) + max(0, cx + r - w + sPx)
                shy = min(0, cy - r) + max(0, cy + r - h + sPx)
                for qy in range(max(0, cy - r - shy), min(h - sPx, cy + r - shy) + 1):
                    for qx in range(max(0, cx - r - shx), min(w - sPx, cx + r - shx) + 1):
                        if (qx, qy) in seen:
                            continue
                        seen.add((qx, qy))
                        cand.append((qx, qy, qt))
                        pats.append(np.concatenate([q[qt:qt + sPt, qy:qy + sPx, qx:qx + sPx].ravel()
                                                    for q in planes]))
            dif = refp[None, :] - np.stack(pats)          # float32
            sq = dif * dif
            d = np.zeros(len(cand), np.float32)
            for j in range(sq.shape[1]):
                d = d + sq[:, j]
            bias = np.array([qx == px and qy == py and (qt == pt or self.mistake != "pred_bias_reference_frame")
                             for (qx, qy, _) in cand])
            d = np.where(bias, d - dsub, d).astype(np.float32)
            sel = partial_sort_order(d, npf)
            traj[qt] = [(cand[i][0], cand[i][1]) for i in sel]
            kept_d += [d[i] for i in sel]
            kept_c += [cand[i] for i in sel]
        kd = np.array(kept_d, np.float32)
        nsim = min(p.nSimilarPatches, len(kd))
        sel = partial_sort_order(kd, nsim)
        tau = np.float32(p.tau) * np.float32(p.tau) * sPx * sPx * sPt
        n = 1
        while 2 * n <= nsim and kd[sel[n]] < tau:
            n *= 2
        return [kept_c[i] for i in sel[:n]]

    # ---- filters: computeBayesEstimateStep{1,2}, nldct.cpp:1799-2242
    def filter1(self, groups, p, B, Bi):
        """groups: list over channels of (N, P) arrays.  Returns estimates and
        the aggregation weight, or None for the equality early-out."""
        N = groups[0].shape[0]
        if N > 1 and all(np.all(g == g[0]) for g in groups):
            return None
        H = haar(N)
        bs2 = float(np.float32(np.float32(p.beta) * np.float32(p.sigma)) ** 2)
        s2 = float(np.float32(p.sigma) * np.float32(p.sigma))
        bms2 = float(np.float32(np.float32(p.betaMean) * np.float32(p.sigma)) ** 2)
        nnz = 0.0
        out = []
        for X in groups:
            Z = H @ (X @ B.T)
            if p.orderInvariance:
                var0 = np.maximum(Z[0] ** 2 - (0 if self.mistake == "oi_no_dc_bias" else s2), 0.0)
                w0 = var0 / (var0 + bms2)
                with np.errstate(invalid="ignore", divide="ignore"):
                    vari = (Z[1:] ** 2).sum(axis=0) / (N - 1)
                vari = np.where(np.isnan(vari), vari, np.maximum(vari - s2, 0.0))
                wi = vari / (vari + bs2)
                Z = np.vstack([Z[:1] * w0, Z[1:] * wi])
                nnz += float((w0 ** 2 + (N - 1) * wi ** 2).sum())
            else:
                near = np.abs(Z ** 2 - bs2) < 2e-5 * bs2
                near[0, 0] = False
                self.ambiguous["threshold"] += int(near.sum())
                keep = Z ** 2 > bs2
                keep[0, 0] = True
                Z = Z * keep
                nnz += float(keep.sum())
            out.append(H.T @ Z @ Bi)
        return out, 1.0 / s2 / nnz

    def filter2(self, gn, gb, p, B):
        N = gn[0].shape[0]
        if N > 1 and all(np.all(g == g[0]) for g in gn):
            return None
        H = haar(N)
        bs2 = float(np.float32(np.float32(p.beta) * np.float32(p.sigma)) ** 2)
        bms2 = float(np.float32(np.float32(p.betaMean) * np.float32(p.sigma)) ** 2)
        s2 = float(np.float32(p.sigma) * np.float32(p.sigma))
        var = 0.0
        out = []
        for Xn, Xb in zip(gn, gb):
            Z = H @ (Xn @ B.T)
            Y = H @ (Xb @ B.T)
            with np.errstate(invalid="ignore", divide="ignore"):
                if p.orderInvariance:
                    v0 = Y[0] ** 2
                    w0 = v0 / (v0 + bms2)
                    vi = (Y[1:] ** 2).sum(axis=0) / (N - 1)
                    wi = vi / (vi + bs2)
                    Z = np.vstack([Z[:1] * w0, Z[1:] * wi])
                    var += float((w0 ** 2 + (N - 1) * wi ** 2).sum())
                else:
                    Wg = Y ** 2 / (Y ** 2 + bs2)
                    Z = Z * Wg
                    var += float((Wg ** 2).sum())
            out.append(H.T @ Z @ B)
        with np.errstate(invalid="ignore", divide="ignore"):
            return out, 1.0 / s2 / var

    # ---- processNlBayes on one tile, nldct.cpp:490-830
    def tile(self, step, noisy, basic, ff, bf, crop, shape):
        p = self.p[step]
        F, Cc, h, w = noisy.shape
        W, Hh, Ff = shape
        sPx, sPt, sW = p.sizePatch, p.sizePatchTime, p.sizeSearchWindow
        sWt = p.sizeSearchTimeRangeFwd + p.sizeSearchTimeRangeBwd + 1
        P = sPx * sPx * sPt
        mask = self.grid(p, crop, (F, h, w), shape, self.mistake)
        grid = set(np.flatnonzero(mask).tolist())
        # bases and window (nldct.cpp:582-737)
        if p.transform == 1:
            B, Bi = (np.asarray(m, np.float64) for m in self.bior)
        else:
            B = np.kron(dct_1d(sPt), np.kron(dct_1d(sPx), dct_1d(sPx)))
            Bi = B
        win = window(sPx, sPt, p.agg_window)
        num = np.zeros((F, Cc, h, w))
        wsum = np.zeros((F, h, w))
        # aggreWeights: one entry per possible neighbour, 1.f at the start of
        # every tile (nldct.cpp:758, 816); a filter sets the first N entries,
        # an equality early-out sets none (nldct.cpp:1935-1939, 2167-2169)
        aggw = np.ones(sW * sW * sWt)
        count = 0
        first_done = False
        for lin in range(F * h * w):
            if not mask[lin]:
                if lin in grid:
                    self.skipped[step] += 1
                continue
            count += 1
            t, rem = divmod(lin, h * w)
            y, x = divmod(rem, w)
            if step == 1:
                nb = self.match(noisy, ff, bf, p, (x, y, t), [0])
            else:
                nb = self.match(basic, ff, bf, p, (x, y, t),
                                [0] if self.mistake == "step2_channel0" else list(range(Cc)))
            self.groups[step].append((crop, (x, y, t), nb))
            N = len(nb)

            def gather(v):
                return [np.stack([v[qt:qt + sPt, c, qy:qy + sPx, qx:qx + sPx].reshape(-1)
                                  for (qx, qy, qt) in nb]) for c in range(Cc)]
            gn = gather(noisy)
            res = self.filter1(gn, p, B, Bi) if step == 1 else \
                self.filter2(gn, gather(basic), p, B)
            if res is None:  # equality early-out: group unchanged, weights stale
                est = gn
                if first_done:
                    self.stale[step] += 1
                if self.mistake == "fresh_weight":
                    aggw[:] = 1.0
            else:
                est, wgt = res
                aggw[:N] = wgt
            first_done = True
            # computeAggregationStep{1,2} (nldct.cpp:2259-2425)
            for n, (qx, qy, qt) in enumerate(nb):
                ww = aggw[n] * win.reshape(sPt, sPx, sPx)
                for c in range(Cc):
                    num[qt:qt + sPt, c, qy:qy + sPx, qx:qx + sPx] += \
                        ww * est[c][n].reshape(sPt, sPx, sPx)
                wsum[qt:qt + sPt, qy:qy + sPx, qx:qx + sPx] += ww
            # mask clear at the first neighbour's 4-argument index
            # t*whc + y*w + x read as a 1-channel position (nldct.cpp:2325-2329)
            qx, qy, qt = nb[0]
            if self.mistake == "clear_reference":
                qx, qy, qt = x, y, t
            k = qt * h * w * Cc + qy * w + qx
            if self.mistake == "clear_one_channel_index":
                k = qt * h * w + qy * w + qx
            if k < F * h * w:
                mask[k] = False
        self.counts[step].append(count)
        # computeWeightedAggregation (nldct.cpp:2436-2448)
        out = np.empty_like(num)
        with np.errstate(invalid="ignore", divide="ignore"):
            for c in range(Cc):
                out[:, c] = np.where(wsum > 0, num[:, c] / np.where(wsum > 0, wsum, 1),
                                     num[:, c] / noisy[:, c])
        return out

    def step(self, step, noisy, basic, ff, bf, nparts):
        p = self.p[step]
        F, Cc, H, W = noisy.shape
        out = np.zeros_like(noisy)
        for crop in sub_divide(W, H, p.boundary, nparts):
            ox, oy, ex, ey, ix0, iy0, ix1, iy1 = crop
            sl = (slice(None), slice(None), slice(oy, ey), slice(ox, ex))
            fsl = None if ff is None else ff[sl]
            bsl = None if bf is None else bf[sl]
            o = self.tile(step, noisy[sl], None if basic is None else basic[sl], fsl, bsl,
                          crop, (W, H, F))
            # subBuildTight: the tile interior (lib_videot.hpp:1102-1163)
            out[:, :, iy0:iy1, ix0:ix1] = o[:, :, iy0 - oy:iy1 - oy, ix0 - ox:ix1 - ox]
        return out

    def bm_all(self, step, image, fflow=None, bflow=None, nparts=2):
        """Neighbour lists at every grid point of every tile (no mask
        clearing), in processing order, as (tile, ref, [global 4-argument
        indices t*whc + y*W + x])."""
        p = self.p[step]
        image = np.asarray(image, np.float64)
        F, Cc, H, W = image.shape
        chans = [0] if step == 1 or self.mistake == "step2_channel0" else list(range(Cc))
        out = []
        for ti, crop in enumerate(sub_divide(W, H, p.boundary, nparts)):
            ox, oy, ex, ey = crop[:4]
            sl = (slice(None), slice(None), slice(oy, ey), slice(ox, ex))
            img = image[sl]
            ff = None if fflow is None else np.asarray(fflow)[sl]
            bf = None if bflow is None else np.asarray(bflow)[sl]
            h, w = ey - oy, ex - ox
            mask = self.grid(p, crop, (F, h, w), (W, H, F), self.mistake)
            for lin in np.flatnonzero(mask):
                t, rem = divmod(int(lin), h * w)
                y, x = divmod(rem, w)
                nb = self.match(img, ff, bf, p, (x, y, t), chans)
                g = [qt * W * H * Cc + (qy + oy) * W + qx + ox for (qx, qy, qt) in nb]
                out.append((ti, t * W * H * Cc + (y + oy) * W + x + ox, g))
        return out

    def run(self, noisy, fflow=None, bflow=None, basic=None, nparts=2):
        """Returns (basic, final, groups_ratio); colour as runNlBayes."""
        noisy = np.asarray(noisy, np.float64)
        F, Cc, H, W = noisy.shape
        p1, p2 = self.p[1], self.p[2]
        ratio = [0.0, 0.0]
        whf = np.float32(W * H * F)
        if p1.sizePatch:
            nz = opponent(noisy, True) if p1.colorSpace == 1 else noisy
            b = self.step(1, nz, None, fflow, bflow, nparts)
            basic = opponent(b, False) if p1.colorSpace == 1 else b
            r = np.float32(0)
            for c in self.counts[1]:
                r = np.float32(r + np.float32(np.float32(100) * np.float32(c)) / whf)
            ratio[0] = float(r)
        final = None
        if p2.sizePatch:
            basic = np.asarray(basic, np.float64)
            nz = opponent(noisy, True) if p2.colorSpace == 1 else noisy
            bz = opponent(basic, True) if p2.colorSpace == 1 else basic
            o = self.step(2, nz, bz, fflow, bflow, nparts)
            final = opponent(o, False) if p2.colorSpace == 1 else o
            r = np.float32(0)
            for c in self.counts[2]:
                r = np.float32(r + np.float32(np.float32(100) * np.float32(c)) / whf)
            ratio[1] = float(r)
        return basic, final, ratio
