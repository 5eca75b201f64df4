"""Integrator parity of the HIP wavefront renderer against the CPU oracle.

Both sides run the same estimator on identical Philox streams.  The integrator is compiled
without FP contraction and both evaluate transcendental functions as the FP64 function
rounded once, so every pixel is expected to agree bit for bit; the bar (tests/_parity.py)
is >= 99.9 % of pixels within 1e-5 relative, and every pixel outside it must be a near-tie
(the oracle's own pixel changes when its closest-hit decisions are made in FP64).
Closed forms (furnace, sky-only scene) are exact.  Ray counts equal the oracle's when the
images agree exactly.

Scenes: Cornell box (compat, at 64x64 and at its full 256x256 x 16 spp size), the sunlit
material mix (compat and pbr: diffuse, GGX metal, mirror, glass, sun NEE), the cutout
scene (textures with negative uv, alpha cut-outs hit by camera and shadow rays, a rotated
16x8 equirect sky), a small terrain with objects (pbr, BVH oracle), and a converged
Cornell image (256x256, 1024 spp) against the FP64 oracle with its mean relative error.
"""
import numpy as np
import pytest

from _parity import assert_parity
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


def _oracle_pair(oracle, spec, res, spp, seed, mode=None, bvh=False):
    """FP32 oracle image and a callable for the same render with FP64 hit decisions."""
    kw = dict(res=res, spp=spp, seed=seed, mode=mode)
    if bvh:
        b = oracle.Bvh(spec.corners())
        ref = oracle.render(spec, trav=oracle.BVH32, bvh=b, **kw)
        return ref, (lambda: oracle.render(spec, trav=oracle.BVH64, bvh=b, **kw)["raw"])
    ref = oracle.render(spec, trav=oracle.BRUTE32, **kw)
    return ref, (lambda: oracle.render(spec, trav=oracle.BRUTE64, **kw)["raw"])


def _check_aovs_and_counts(gpu, ref, c, shadow=True):
    for k in ("albedo", "normal", "depth"):
        ok = (gpu[k] == ref[k]).all(axis=-1).mean()
        assert ok >= 0.999, (k, ok)
    st, rs = gpu["stats"], ref["stats"]
    if c["outside"] == 0:
        assert st["extension_rays"] == rs["ext_rays"]
        assert st["total_rays"] == rs["ref_segments"]
        if shadow:
            assert st["shadow_rays"] == rs["shadow_rays"]
    else:
        assert abs(st["extension_rays"] - rs["ext_rays"]) <= 1e-3 * rs["ext_rays"]


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
    print(f"cornell 64x64 width {width}", flush=True)
    gpu = _render(spec, res, spp, seed=seed, width=width)
    ref, alt = _oracle_pair(oracle, spec, res, spp, seed)
    c = assert_parity(gpu["raw"], ref["raw"], alt)
    _check_aovs_and_counts(gpu, ref, c, shadow=False)
    np.testing.assert_allclose(gpu["progress"][..., :3], oracle.display(gpu["raw"], spp), rtol=1e-5, atol=1e-6)


def test_cornell_full_size_matches_oracle(oracle):
    # the Cornell configuration at its own size: 256x256, 16 spp, depth 8, RR after bounce 3
    spec = scenes.cornell(1)
    res, spp, seed = spec.res, spec.spp, 1
    print(f"cornell {res} x {spp} spp", flush=True)
    gpu = _render(spec, res, spp, seed=seed)
    ref, alt = _oracle_pair(oracle, spec, res, spp, seed)
    c = assert_parity(gpu["raw"], ref["raw"], alt)
    _check_aovs_and_counts(gpu, ref, c, shadow=False)


@pytest.mark.parametrize("mode", [0, 1])
def test_sunlit_materials_match_oracle(oracle, mode):
    spec = scenes.sunlit(4, mode=mode)
    res, spp, seed = spec.res, 4, 21
    print(f"sunlit mode {mode}", flush=True)
    gpu = _render(spec, res, spp, seed=seed)
    ref, alt = _oracle_pair(oracle, spec, res, spp, seed)
    c = assert_parity(gpu["raw"], ref["raw"], alt)
    _check_aovs_and_counts(gpu, ref, c)


@pytest.mark.parametrize("mode", [0, 1])
def test_cutout_textures_alpha_sky_match_oracle(oracle, mode):
    spec = scenes.cutout(5, mode=mode)
    res, spp, seed = spec.res, 4, 3
    print(f"cutout mode {mode}", flush=True)
    gpu = _render(spec, res, spp, seed=seed)
    ref, alt = _oracle_pair(oracle, spec, res, spp, seed)
    c = assert_parity(gpu["raw"], ref["raw"], alt)
    _check_aovs_and_counts(gpu, ref, c)
    # the branches under test ran: sky seen, shadow rays continued through cut-outs
    assert gpu["stats"]["shadow_rays"] > 0
    assert ref["stats"]["shadow_rays"] > 0


def test_terrain_pbr_matches_oracle(oracle):
    # the terrain-with-objects workload at a size the FP32 CPU BVH oracle renders in seconds:
    # fBm terrain, icospheres of every material, sun NEE + sky, pbr estimator, RR from 3
    spec = scenes.terrain(grid=96, spheres=150, seed=3, res=(64, 36))
    res, spp, seed = (64, 36), 4, 13
    print("terrain 96 / 150", flush=True)
    gpu = _render(spec, res, spp, seed=seed)
    ref, alt = _oracle_pair(oracle, spec, res, spp, seed, bvh=True)
    c = assert_parity(gpu["raw"], ref["raw"], alt)
    _check_aovs_and_counts(gpu, ref, c)


def test_cornell_converged_vs_fp64_oracle(oracle):
    """The converged image against an FP64 CPU render of the same paths: Cornell box at
    256x256, 1024 spp.  The mean relative error, sum |gpu - fp64| / sum fp64 over all
    pixels and channels, stays below 2e-5."""
    spec = scenes.cornell(1)
    res, spp, seed = (256, 256), 1024, 17
    print("cornell converged 256x256 x 1024 spp", flush=True)
    gpu = _render(spec, res, spp, seed=seed)
    ref = oracle.render(spec, res=res, spp=spp, seed=seed, precision=64, trav=oracle.BRUTE64, aovs=False)["raw"]
    g = gpu["raw"].astype(np.float64) / spp
    r = ref / spp
    mre = float(np.abs(g - r).sum() / r.sum())
    px = np.abs(g - r).sum(axis=-1) / np.maximum(r.sum(axis=-1), 1e-6)
    print(f"converged mean relative error {mre:.3e}, pixel p99 {np.percentile(px, 99):.3e}, "
          f"max {px.max():.3e}", flush=True)
    assert mre <= 2e-5
