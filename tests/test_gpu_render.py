"""Integrator parity of the HIP wavefront renderer against the CPU oracle.

Both sides run the same estimator on identical Philox streams, so the per-pixel results
agree up to FP32 evaluation order (FMA contraction and transcendental rounding on the
GPU) and the rare path that such a difference sends to another triangle.  Tolerances:
  * closed forms (furnace, sky-only scene): exact;
  * compat / pbr images vs the FP32 oracle: >= 99% of pixels within 1e-3 relative
    (+1e-4 absolute), image mean within 0.5%;
  * ray counts within 0.5% (paths that diverge trace a different number of rays).
"""
import numpy as np
import pytest

from crender_amd import Context, Renderer, scene_from_spec, scenes

pytestmark = pytest.mark.gpu


def _render(spec, res, spp, seed=0, mode=None, rr=None, bounces=None, width=0):
    ctx = Context(0, bvh_width=width)
    sc = scene_from_spec(ctx, spec)
    W, H = res
    r = Renderer(ctx, sc, W, H, spec.max_bounces if bounces is None else bounces)
    r.set_resolution(W, H)
    r.set_options(mode=spec.mode if mode is None else mode, rr_start=spec.rr_start if rr is None else rr, seed=seed)
    r.render(spp)
    out = {"raw": r.raw(), "albedo": r.current_albedos(), "normal": r.current_normals(),
           "depth": r.current_depths(), "progress": r.current_progress(), "stats": r.current_stats(),
           "n": r.current_sample_count()}
    r.close()
    sc.close()
    ctx.close()
    return out


def _agree(gpu, ref, rtol=1e-3, atol=1e-4):
    return float(np.isclose(gpu, ref, rtol=rtol, atol=atol).all(axis=-1).mean())


def test_furnace_closed_form():
    spec = scenes.furnace(albedo=0.5, emission=1.0)
    out = _render(spec, (32, 32), spp=3, bounces=8)
    expected = np.float32(sum(0.5 ** k for k in range(1, 9)))
    assert np.all(out["raw"] == np.float32(3) * expected)
    assert out["n"] == 3
    st = out["stats"]
    assert st["extension_rays"] == 3 * 32 * 32 * 8
    assert st["total_rays"] == 3 * 32 * 32 * 9      # total_bounces = 1 + 8 per path


def test_sky_only_scene_is_the_sky():
    spec = scenes.soup(10)
    spec.vertices[:] = 100.0        # geometry far behind the camera's view
    spec.camera = scenes.camera_dict((0.0, 0.0, 0.0), rotation=(180.0, 0.0, 0.0), fov=30.0)
    spec.sky = np.full((1, 1, 4), 0.25, np.float32)
    out = _render(spec, (16, 16), spp=2, bounces=4)
    assert np.all(out["raw"] == np.float32(0.5))
    assert np.all(out["albedo"][..., :3] == np.float32(0.25))
    assert np.all(out["depth"][..., :3] == 0.0)
    # progress = pow(clamp(sum / n, 0, 1), 1 / 2.2)
    np.testing.assert_allclose(out["progress"][..., :3], np.float32(0.25) ** (1 / 2.2), rtol=1e-6)


@pytest.mark.parametrize("width", [4, 8])
def test_cornell_compat_matches_oracle(oracle, width):
    spec = scenes.cornell(1)
    res, spp, seed = (64, 64), 4, 9
    gpu = _render(spec, res, spp, seed=seed, width=width)
    ref = oracle.render(spec, res=res, spp=spp, seed=seed, trav=oracle.BRUTE32)
    assert _agree(gpu["raw"], ref["raw"]) >= 0.99
    assert abs(gpu["raw"].mean() - ref["raw"].mean()) <= 5e-3 * ref["raw"].mean()
    # AOVs of the last pass
    assert _agree(gpu["albedo"], ref["albedo"], rtol=1e-5, atol=1e-6) >= 0.99
    assert _agree(gpu["normal"], ref["normal"], rtol=1e-5, atol=1e-5) >= 0.99
    assert _agree(gpu["depth"], ref["depth"], rtol=1e-4, atol=1e-6) >= 0.99
    st, rs = gpu["stats"], ref["stats"]
    assert abs(st["extension_rays"] - rs["ext_rays"]) <= 5e-3 * rs["ext_rays"]
    assert abs(st["total_rays"] - rs["ref_segments"]) <= 5e-3 * rs["ref_segments"]
    np.testing.assert_allclose(gpu["progress"][..., :3], oracle.display(gpu["raw"], spp), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("mode", [0, 1])
def test_sunlit_materials_match_oracle(oracle, mode):
    spec = scenes.sunlit(4, mode=mode)
    res, spp, seed = spec.res, 4, 21
    gpu = _render(spec, res, spp, seed=seed)
    ref = oracle.render(spec, res=res, spp=spp, seed=seed, trav=oracle.BRUTE32)
    assert _agree(gpu["raw"], ref["raw"], rtol=2e-3, atol=1e-3) >= 0.98
    assert abs(gpu["raw"].mean() - ref["raw"].mean()) <= 1e-2 * ref["raw"].mean()
    st, rs = gpu["stats"], ref["stats"]
    assert abs(st["extension_rays"] - rs["ext_rays"]) <= 1e-2 * rs["ext_rays"]
    if mode == 0:
        assert abs(st["shadow_rays"] - rs["shadow_rays"]) <= 1e-2 * rs["shadow_rays"]


def test_terrain_pbr_matches_oracle(oracle):
    # the terrain-with-objects workload at a size the FP32 CPU BVH oracle renders in seconds:
    # fBm terrain, icospheres of every material, sun NEE + sky, pbr estimator, RR from 3
    spec = scenes.terrain(grid=96, spheres=150, seed=3, res=(64, 36))
    res, spp, seed = (64, 36), 4, 13
    gpu = _render(spec, res, spp, seed=seed)
    bvh = oracle.Bvh(spec.corners())
    ref = oracle.render(spec, res=res, spp=spp, seed=seed, trav=oracle.BVH32, bvh=bvh)
    assert _agree(gpu["raw"], ref["raw"], rtol=2e-3, atol=1e-3) >= 0.98
    assert abs(gpu["raw"].mean() - ref["raw"].mean()) <= 1e-2 * ref["raw"].mean()
    st, rs = gpu["stats"], ref["stats"]
    assert abs(st["extension_rays"] - rs["ext_rays"]) <= 1e-2 * rs["ext_rays"]
