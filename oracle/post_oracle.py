"""CPU restatement of the renderer's image post-processing, for tests.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of crt_post_process and
crt_denoise, never by the product.

post_process():  cr::post_processor::process (src/render/post/post_processor.cpp:110-248)
  with _brightness (post_processor.cpp:302-311), _blur (post_processor.cpp:250-300,
  assets/app/shaders/blur.comp) and the composite shader
  (assets/app/shaders/post_process.comp:36-134), in float32 with the shaders' order of
  operations and no fused multiply-add.  Deviations, both where the reference is undefined:
  texels outside the image read as 0 (texelFetch out of range), and the brightness pass
  indexes rows by width (post_processor.cpp:304 divides by height, which reads out of
  bounds for non-square images).
atrous():  the edge-avoiding a-trous filter of crt_denoise (Dammertz et al. 2010); it
  restates this library's own denoiser (the reference calls Open Image Denoise).
"""
from __future__ import annotations

import numpy as np

F = np.float32
BLUR_W = np.array([0.19744746769063704, 0.1746973469158936, 0.12099884565428047, 0.06560233156931679,
                   0.027839605612666265, 0.009246250740395456, 0.002403157286908872, 0.00048872837522002],
                  np.float32)   # blur.comp:11-19


def brightness(img, threshold):
    """post_processor.cpp:302-311: texels with luma below the threshold become (0,0,0,1)."""
    lum = (img[..., 0] * F(0.2126) + img[..., 1] * F(0.7152)) + img[..., 2] * F(0.0722)
    out = img.copy()
    out[lum < F(threshold)] = np.array([0, 0, 0, 1], np.float32)
    return out


def _shift(a, k, axis):
    """a[x + k] along axis with zeros outside."""
    out = np.zeros_like(a)
    n = a.shape[axis]
    if abs(k) >= n:
        return out
    if axis == 1:
        if k > 0:
            out[:, :n - k] = a[:, k:]
        else:
            out[:, -k:] = a[:, :n + k]
    else:
        if k > 0:
            out[:n - k] = a[k:]
        else:
            out[-k:] = a[:n + k]
    return out


def blur_pass(src, horizontal, cover):
    """blur.comp:21-44 for the pixels the 8x8 dispatch covers; the rest stay cleared (0)."""
    axis = 1 if horizontal else 0
    c = src[..., :3]
    r = c * BLUR_W[0]
    for k in range(1, 8):
        r = r + _shift(c, k, axis) * BLUR_W[k]
        r = r + _shift(c, -k, axis) * BLUR_W[k]
    out = np.zeros_like(src)
    cw, ch = cover
    out[:ch, :cw, :3] = r[:ch, :cw]
    out[:ch, :cw, 3] = 1.0
    return out


def bloom_image(img, threshold, cover):
    """post_processor.cpp:117-120,250-300: brightness, then 10 ping-pong blur passes."""
    t0 = brightness(img, threshold)
    t1 = np.zeros_like(t0)
    for i in range(10):
        if i % 2 == 0:
            t1 = blur_pass(t0, True, cover)
        else:
            t0 = blur_pass(t1, False, cover)
    return t0


def _uncharted_map(x):
    A, B, C, D, E, Fv = F(0.15), F(0.50), F(0.10), F(0.20), F(0.02), F(0.30)
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * Fv)) - E / Fv


def post_process(img, bloom=False, threshold=0.7, strength=1.0, gray=False, tonemap=False, tonemap_type=0,
                 exposure=1.0, gamma=2.2, full_coverage=False):
    """post_process.comp:116-134 over an (H, W, 4) float32 image."""
    img = np.asarray(img, np.float32)
    if not (bloom or gray or tonemap):
        return img.copy()                       # post_processor.cpp:112-114
    H, W = img.shape[:2]
    cover = (W, H) if full_coverage else (W // 8 * 8, H // 8 * 8)
    c = img[..., :3].copy()
    if bloom:
        c = c + bloom_image(img, threshold, cover)[..., :3] * F(strength)
    if gray:
        g = (F(0.2126) * c[..., 0] + F(0.7162) * c[..., 1]) + F(0.0722) * c[..., 2]
        c = np.stack([g, g, g], -1)
    if tonemap:
        ig = F(1.0) / F(gamma)
        c = c * F(exposure)
        with np.errstate(invalid="ignore", divide="ignore"):
            if tonemap_type == 0:
                c = np.power(c, ig)
            elif tonemap_type == 1:
                c = np.power(c / (c + F(1.0)), ig)
            elif tonemap_type == 2:
                v = np.maximum(F(0.0), c - F(0.004))
                c = (v * (F(6.2) * v + F(0.5))) / (v * (F(6.2) * v + F(1.7)) + F(0.06))
            elif tonemap_type == 3:
                ws = F(1.0) / _uncharted_map(F(11.2))
                c = np.power(_uncharted_map(F(2.0) * c) * ws, ig)
    out = np.zeros_like(img)
    cw, ch = cover
    out[:ch, :cw, :3] = c[:ch, :cw]
    out[:ch, :cw, 3] = 1.0
    return out


def atrous(colour, normal, albedo, levels=5, sc=2.0, sn=0.2, sa=0.2):
    """Edge-avoiding a-trous wavelet filter: per level, a 5x5 B3-spline kernel with holes of
    2^level pixels, each tap weighted by exp(-(|dc|^2/sc_l^2 + |dn|^2/sn^2 + |da|^2/sa^2)),
    sc_l = sc * 2^-level; float64."""
    c = np.asarray(colour, np.float64)[..., :3]
    n = np.asarray(normal, np.float64)[..., :3]
    a = np.asarray(albedo, np.float64)[..., :3]
    H, W = c.shape[:2]
    kern = np.array([1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16])
    for lvl in range(levels):
        step = 1 << lvl
        scl = sc * 2.0 ** -lvl
        acc = np.zeros_like(c)
        wsum = np.zeros((H, W))
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                ys, xs = np.arange(H) + dy * step, np.arange(W) + dx * step
                vy, vx = (ys >= 0) & (ys < H), (xs >= 0) & (xs < W)
                valid = vy[:, None] & vx[None, :]
                yy, xx = np.clip(ys, 0, H - 1), np.clip(xs, 0, W - 1)
                cq, nq, aq = c[yy][:, xx], n[yy][:, xx], a[yy][:, xx]
                d = (np.sum((cq - c) ** 2, -1) / scl ** 2 + np.sum((nq - n) ** 2, -1) / sn ** 2 +
                     np.sum((aq - a) ** 2, -1) / sa ** 2)
                w = kern[dx + 2] * kern[dy + 2] * np.exp(-d) * valid
                acc += w[..., None] * cq
                wsum += w
        c = acc / wsum[..., None]
    out = np.ones((H, W, 4))
    out[..., :3] = c
    return out
