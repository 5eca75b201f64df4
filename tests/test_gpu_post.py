"""GPU post-processing and denoising against the CPU restatements (oracle/post_oracle.py).

  * post-processing without a pow() (bloom, gray scale, Hejl-Burgess-Dawson) is bit-exact:
    both sides evaluate the shaders' float32 expressions in the same order without fused
    multiply-add; with pow() (linear / Reinhard / Uncharted 2) within 2e-6 relative;
  * the denoiser matches the float64 restatement of its filter within 1e-4, keeps constant
    images and lowers the error of a noisy render against a converged one.
"""
import numpy as np
import pytest

import post_oracle as P
from crender_amd import Context, Renderer, denoise, post_process, scene_from_spec, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _image(seed, h=45, w=70):
    rng = np.random.default_rng(seed)
    img = rng.random((h, w, 4), dtype=np.float32)
    img[..., :3] *= 1.5
    img[10:14, 20:30, :3] = 4.0       # bright block for the bloom
    return img


@pytest.mark.parametrize("opts", [
    dict(bloom=True), dict(bloom=True, threshold=0.4, strength=0.5), dict(gray=True),
    dict(tonemap=True, tonemap_type=2, exposure=1.3), dict(bloom=True, gray=True, tonemap=True, tonemap_type=2),
    dict(bloom=True, full_coverage=True),
])
def test_post_exact(ctx, opts):
    img = _image(1)
    print("post", opts, flush=True)
    got = post_process(ctx, img, **opts)
    want = P.post_process(img, **opts)
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize("tt", [0, 1, 3])
def test_post_tonemap_pow(ctx, tt):
    img = _image(2)
    got = post_process(ctx, img, bloom=True, tonemap=True, tonemap_type=tt, exposure=0.8, gamma=2.2)
    want = P.post_process(img, bloom=True, tonemap=True, tonemap_type=tt, exposure=0.8, gamma=2.2)
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-7)


def test_post_passthrough(ctx):
    img = _image(3)
    assert np.array_equal(post_process(ctx, img), img)


def test_denoise_matches_restatement(ctx):
    rng = np.random.default_rng(4)
    c = rng.random((40, 56, 4), dtype=np.float32)
    n = rng.random((40, 56, 4), dtype=np.float32)
    a = rng.random((40, 56, 4), dtype=np.float32)
    got = denoise(ctx, c, n, a)
    want = P.atrous(c, n, a)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-5)
    const = np.full((40, 56, 4), 0.3, np.float32)
    np.testing.assert_allclose(denoise(ctx, const, n, a)[..., :3], 0.3, rtol=1e-6)


def test_denoise_lowers_render_error(ctx):
    # cr::denoise is applied to current_progress() (ui.h:607-611): the clamped, gamma-encoded
    # display image, guided by the last pass's normals and albedos
    spec = scenes.cornell(1)
    sc = scene_from_spec(ctx, spec)
    res = (96, 96)
    imgs = {}
    for spp in (16, 1024):
        r = Renderer(ctx, sc, res[0], res[1], 8)
        r.set_resolution(*res)
        r.set_options(mode=1, rr_start=3, seed=spp)
        r.render(spp)
        imgs[spp] = (r.current_progress(), r.current_normals(), r.current_albedos())
        r.close()
    noisy, nrm, alb = imgs[16]
    ref = imgs[1024][0][..., :3]
    out = denoise(ctx, noisy, nrm, alb)
    e_noisy = np.mean((noisy[..., :3] - ref) ** 2)
    e_den = np.mean((out[..., :3] - ref) ** 2)
    print(f"denoise: noisy MSE {e_noisy:.4e}, denoised {e_den:.4e}", flush=True)
    assert e_den < 0.5 * e_noisy
    sc.close()
