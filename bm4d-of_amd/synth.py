"""Seeded procedural test scenes and noise (numpy).

The reference ships no sample data and seeds its noise tool with time and
pid (src/utils/lib_videot.hpp:491-492), so inputs are generated here from
explicit seeds:

* image(): shapes painted over 0 plus oriented sinusoids added globally
  (single grayscale frame);
* video(): textured ellipses / rectangles, each under its own per-frame
  affine motion, rendered by bilinear inverse warp over a smooth background
  gradient;
* add_noise(): i.i.d. Gaussian noise, unclipped, in float32.

Layout of every returned video is the reference's (F, C, H, W) float32.
"""
import numpy as np


def _shape_mask(xx, yy, cx, cy, ax, ay, ang, kind):
    c, s = np.cos(ang), np.sin(ang)
    u = (xx - cx) * c + (yy - cy) * s
    v = -(xx - cx) * s + (yy - cy) * c
    if kind == 0:
        return (u / ax) ** 2 + (v / ay) ** 2 <= 1.0
    return (np.abs(u) <= ax) & (np.abs(v) <= ay)


def image(width=256, height=256, n_shapes=12, n_sines=3, seed=100):
    """Still grayscale scene, (1, 1, H, W) float32."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    img = np.zeros((height, width))
    for _ in range(n_shapes):
        cx, cy = rng.uniform(0, width), rng.uniform(0, height)
        ax, ay = rng.uniform(8, 64), rng.uniform(8, 64)
        ang = rng.uniform(0, np.pi)
        val = rng.uniform(0, 255)
        kind = int(rng.integers(0, 2))
        img[_shape_mask(xx, yy, cx, cy, ax, ay, ang, kind)] = val
    for _ in range(n_sines):
        period = rng.uniform(4, 16)
        ang = rng.uniform(0, np.pi)
        phase = rng.uniform(0, 2 * np.pi)
        img += 20.0 * np.sin(2 * np.pi * (xx * np.cos(ang) + yy * np.sin(ang)) / period + phase)
    return img.astype(np.float32)[None, None]


def _bilinear(layer, x, y):
    """Sample layer (H, W, K) at float coords with zero outside."""
    H, W = layer.shape[:2]
    x0 = np.floor(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    fx = (x - x0)[..., None]
    fy = (y - y0)[..., None]
    out = np.zeros(x.shape + (layer.shape[2],))
    for dy in (0, 1):
        for dx in (0, 1):
            xi, yi = x0 + dx, y0 + dy
            ok = (xi >= 0) & (xi < W) & (yi >= 0) & (yi < H)
            wgt = (fx if dx else 1 - fx) * (fy if dy else 1 - fy)
            vals = np.zeros_like(out)
            vals[ok] = layer[yi[ok], xi[ok]]
            out += wgt * vals
    return out


def video(width=960, height=540, frames=20, channels=1, n_shapes=20, seed=101,
          max_shift=3.0, max_rot_deg=1.0, threads=1):
    """Moving textured scene, (F, C, H, W) float32 in [0, 255]-ish range.
    Frames are rendered independently; `threads` > 1 renders them on a
    thread pool with the same result."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    # smooth background gradient per channel
    g0 = rng.uniform(40, 200, size=channels)
    gx = rng.uniform(-60, 60, size=channels)
    gy = rng.uniform(-60, 60, size=channels)
    bg = np.stack([g0[c] + gx[c] * (xx / width - 0.5) + gy[c] * (yy / height - 0.5)
                   for c in range(channels)], axis=-1)
    shapes = []
    for _ in range(n_shapes):
        sz = rng.uniform(12, max(16.0, min(width, height) / 5))
        ax, ay = sz * rng.uniform(0.6, 1.4), sz * rng.uniform(0.6, 1.4)
        kind = int(rng.integers(0, 2))
        base = rng.uniform(0, 255, size=channels)
        period = rng.uniform(4, 16, size=channels)
        tang = rng.uniform(0, np.pi, size=channels)
        amp = rng.uniform(10, 30, size=channels)
        # layer in shape-local coordinates, centred in a (2R+1)^2 box
        R = int(np.ceil(max(ax, ay))) + 2
        ly, lx = np.mgrid[-R:R + 1, -R:R + 1].astype(np.float64)
        m = _shape_mask(lx, ly, 0.0, 0.0, ax, ay, 0.0, kind)
        tex = np.stack([base[c] + amp[c] * np.sin(2 * np.pi * (lx * np.cos(tang[c]) + ly *
                        np.sin(tang[c])) / period[c]) for c in range(channels)], axis=-1)
        layer = np.concatenate([tex * m[..., None], m[..., None].astype(np.float64)], axis=-1)
        shapes.append(dict(
            R=R, layer=layer,
            cx=rng.uniform(0, width), cy=rng.uniform(0, height), ang=rng.uniform(0, np.pi),
            vx=rng.uniform(-max_shift, max_shift), vy=rng.uniform(-max_shift, max_shift),
            va=np.deg2rad(rng.uniform(-max_rot_deg, max_rot_deg))))
    out = np.zeros((frames, channels, height, width), np.float32)

    def render(t):
        img = bg.copy()
        for s in shapes:
            cx, cy = s["cx"] + s["vx"] * t, s["cy"] + s["vy"] * t
            ang = s["ang"] + s["va"] * t
            R = s["R"]
            # bounding box of the rotated layer
            x0, x1 = int(max(0, np.floor(cx - 1.5 * R))), int(min(width, np.ceil(cx + 1.5 * R) + 1))
            y0, y1 = int(max(0, np.floor(cy - 1.5 * R))), int(min(height, np.ceil(cy + 1.5 * R) + 1))
            if x1 <= x0 or y1 <= y0:
                continue
            sx, sy = xx[y0:y1, x0:x1] - cx, yy[y0:y1, x0:x1] - cy
            c, sn = np.cos(ang), np.sin(ang)
            u = sx * c + sy * sn + R
            v = -sx * sn + sy * c + R
            smp = _bilinear(s["layer"], u, v)
            a = smp[..., -1:]
            img[y0:y1, x0:x1] = img[y0:y1, x0:x1] * (1 - a) + smp[..., :-1]
        out[t] = np.moveaxis(img, -1, 0).astype(np.float32)

    if threads > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(render, range(frames)))
    else:
        for t in range(frames):
            render(t)
    return out


def add_noise(clean, sigma, seed):
    """clean + N(0, sigma^2) i.i.d., unclipped, float32."""
    rng = np.random.default_rng(seed)
    return (clean + rng.normal(0.0, sigma, size=clean.shape)).astype(np.float32)
