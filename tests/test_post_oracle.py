"""Pins the CPU restatement of the post-processor (oracle/post_oracle.py) against closed
forms of blur.comp / post_process.comp and the dispatch coverage of post_processor.cpp."""
import numpy as np

import post_oracle as P

F = np.float32


def test_blur_impulse_response_is_the_kernel():
    img = np.zeros((16, 40, 4), np.float32)
    img[8, 20, :3] = 1.0
    out = P.blur_pass(img, True, (40, 16))
    row = out[8, :, 0]
    np.testing.assert_array_equal(row[20], P.BLUR_W[0])
    for k in range(1, 8):
        assert row[20 + k] == P.BLUR_W[k] and row[20 - k] == P.BLUR_W[k]
    assert row[:13].sum() == 0 and row[28:].sum() == 0
    assert np.all(out[np.arange(16) != 8, :, 0] == 0)      # horizontal pass: other rows untouched
    assert np.all(out[..., 3] == 1.0)
    v = P.blur_pass(img, False, (40, 16))[:, 20, 1]
    for k in range(1, 8):
        assert v[8 + k] == P.BLUR_W[k] and v[8 - k] == P.BLUR_W[k]


def test_blur_kernel_sums_to_one_and_edges_read_zero():
    total = P.BLUR_W[0] + 2 * P.BLUR_W[1:].sum(dtype=np.float64)
    assert abs(total - 1.0) < 1e-6
    img = np.ones((8, 24, 4), np.float32)
    out = P.blur_pass(img, True, (24, 8))
    assert abs(out[4, 12, 0] - total) < 1e-6                    # interior
    assert abs(out[4, 0, 0] - (P.BLUR_W[0] + P.BLUR_W[1:].sum(dtype=np.float64))) < 1e-6   # left edge: half kernel


def test_dispatch_coverage_leaves_the_remainder_black():
    img = np.full((13, 21, 4), 0.5, np.float32)
    out = P.post_process(img, gray=True)
    assert np.all(out[:8, :16, 3] == 1.0)
    assert np.all(out[8:] == 0) and np.all(out[:, 16:] == 0)
    full = P.post_process(img, gray=True, full_coverage=True)
    assert np.all(full[..., 3] == 1.0)


def test_tonemapping_closed_forms():
    x = np.linspace(0.0, 3.0, 49, dtype=np.float32)
    img = np.zeros((7, 8, 4), np.float32)
    img[..., 0] = x.reshape(7, 7)[:, :1].repeat(8, 1)
    img[..., 1] = 0.25
    img[..., 2] = 1.0
    # linear, gamma 1: colour * exposure
    out = P.post_process(img, tonemap=True, tonemap_type=0, exposure=2.0, gamma=1.0, full_coverage=True)
    np.testing.assert_allclose(out[..., :3], img[..., :3] * 2.0, rtol=1e-7)
    # Reinhard, gamma 1: e x / (e x + 1)
    out = P.post_process(img, tonemap=True, tonemap_type=1, exposure=2.0, gamma=1.0, full_coverage=True)
    e = img[..., :3].astype(np.float64) * 2.0
    np.testing.assert_allclose(out[..., :3], e / (e + 1.0), rtol=1e-6)
    # Hejl-Burgess-Dawson: 0 at x <= 0.004, tends to 6.2/6.2 = 1
    tiny = np.zeros((8, 8, 4), np.float32)
    tiny[..., :3] = 0.003
    assert np.all(P.post_process(tiny, tonemap=True, tonemap_type=2)[..., :3] == 0.0)
    big = np.full((8, 8, 4), 1e4, np.float32)
    np.testing.assert_allclose(P.post_process(big, tonemap=True, tonemap_type=2)[..., :3], 1.0, rtol=1e-3)
    # Uncharted 2: white point 11.2 = 2 * 5.6 maps to 1
    w = np.full((8, 8, 4), 5.6, np.float32)
    np.testing.assert_allclose(P.post_process(w, tonemap=True, tonemap_type=3, gamma=1.0)[..., :3], 1.0, rtol=1e-6)


def test_gray_weights_and_brightness_threshold():
    img = np.zeros((8, 8, 4), np.float32)
    img[..., 0], img[..., 1], img[..., 2] = 1.0, 0.0, 0.0
    np.testing.assert_allclose(P.post_process(img, gray=True)[..., :3], 0.2126, rtol=1e-7)
    img[..., 0], img[..., 1] = 0.0, 1.0
    np.testing.assert_allclose(P.post_process(img, gray=True)[..., :3], 0.7162, rtol=1e-7)   # post_process.comp's weight
    b = np.zeros((2, 2, 4), np.float32)
    b[0, 0] = [1, 1, 1, 0.3]      # luma 1 >= 0.7: kept
    b[0, 1] = [0.5, 0.5, 0.5, 0.3]  # luma 0.5 < 0.7: black, alpha 1
    out = P.brightness(b, 0.7)
    assert out[0, 0].tolist() == [1, 1, 1, np.float32(0.3)] and out[0, 1].tolist() == [0, 0, 0, 1]


def test_atrous_keeps_constants_and_smooths_noise():
    rng = np.random.default_rng(0)
    const = np.full((24, 32, 4), 0.4)
    n = np.full((24, 32, 4), 0.5)
    out = P.atrous(const, n, n)
    np.testing.assert_allclose(out[..., :3], 0.4, rtol=1e-12)
    noisy = const.copy()
    noisy[..., :3] += rng.normal(0, 0.05, (24, 32, 3))
    out = P.atrous(noisy, n, n)
    assert np.mean((out[..., :3] - 0.4) ** 2) < 0.1 * np.mean((noisy[..., :3] - 0.4) ** 2)


def test_atrous_without_edge_stopping_is_the_b3_spline_cascade():
    """With every edge-stopping width huge, each level is the plain 5x5 B3-spline kernel
    (1, 4, 6, 4, 1) / 16 with holes of 2^level pixels: away from the border (62 pixels: the
    sum of the levels' radii) the filter equals that cascade of convolutions computed by
    scipy, and at the border a tap outside the image drops out with its weight renormalised
    (a corner pixel of one level: the taps at offsets 0..2 in both directions)."""
    from scipy.ndimage import convolve

    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1, (150, 140, 4))
    flat = np.full(img.shape, 0.5)
    out = P.atrous(img, flat, flat, levels=5, sc=1e9, sn=1e9, sa=1e9)[..., :3]
    k1 = np.array([1, 4, 6, 4, 1], np.float64) / 16.0
    want = img[..., :3].copy()
    for lvl in range(5):
        k = np.zeros(4 * (1 << lvl) + 1)
        k[:: 1 << lvl] = k1
        k2 = np.outer(k, k)
        want = np.stack([convolve(want[..., ch], k2, mode="constant") for ch in range(3)], -1)
    np.testing.assert_allclose(out[62:-62, 62:-62], want[62:-62, 62:-62], rtol=0, atol=1e-12)
    one = P.atrous(img, flat, flat, levels=1, sc=1e9, sn=1e9, sa=1e9)[0, 0, :3]
    w = np.outer(k1[2:], k1[2:])
    np.testing.assert_allclose(one, np.einsum("yx,yxc->c", w, img[:3, :3, :3]) / w.sum(), rtol=1e-13)


def test_atrous_with_sharp_normal_edges_is_the_identity():
    """With a tiny normal width and a different normal at every pixel, every tap but the
    pixel itself gets weight exp(-huge) = 0, so the output is the input ((w c) / w: within
    an ulp)."""
    rng = np.random.default_rng(4)
    img = rng.uniform(0, 1, (20, 30, 4))
    nrm = rng.normal(0, 1, (20, 30, 4))
    nrm[..., :3] /= np.linalg.norm(nrm[..., :3], axis=-1, keepdims=True)
    alb = np.full(img.shape, 0.5)
    out = P.atrous(img, nrm, alb, levels=5, sn=1e-4)
    np.testing.assert_allclose(out[..., :3], img[..., :3], rtol=3e-16, atol=0)


def test_atrous_defaults_halve_display_error_of_a_16spp_render():
    """The default tolerances on the renderer's own output: a 16 spp Cornell box display
    image, guided by that render's normals and albedos, against a 1024 spp one."""
    import pyoracle as po
    from crender_amd import scenes

    spec = scenes.cornell(1)
    bvh = po.Bvh(np.ascontiguousarray(spec.corners(), dtype=np.float32))
    kw = dict(res=(96, 96), mode=1, rr_start=3, max_bounces=8, trav=po.BVH32, bvh=bvh, nthreads=8)
    a = po.render(spec, spp=16, seed=16, **kw)
    b = po.render(spec, spp=1024, seed=1024, aovs=False, **kw)
    noisy, ref = po.display(a["raw"], 16), po.display(b["raw"], 1024)
    out = P.atrous(noisy, a["normal"], a["albedo"])[..., :3]
    assert np.mean((out - ref) ** 2) < 0.5 * np.mean((noisy - ref) ** 2)
