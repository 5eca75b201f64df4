"""Renderer / scene API behaviour on the GPU (the cr::renderer and cr::scene contract).

  * start() clears the accumulation and renders in the background until the target
    spp; pause() waits for the running pass (renderer.cpp:116-183);
  * checkpoint + resume equals an uninterrupted run bit for bit (counter-based RNG);
  * background and synchronous passes produce the same image;
  * several models: geomID / per-model primID (scene.cpp:83-95);
  * error paths return codes instead of exiting;
  * results against the committed golden fixtures.
"""
from pathlib import Path

import numpy as np
import pytest

from crender_amd import Context, CrtError, Renderer, Scene, scene_from_spec, scenes

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
MISS = 0xFFFFFFFF


@pytest.fixture(scope="module")
def cornell_scene():
    ctx = Context(0)
    sc = scene_from_spec(ctx, scenes.cornell(1))
    yield ctx, sc
    sc.close()
    ctx.close()


def test_background_passes_equal_synchronous(cornell_scene):
    ctx, sc = cornell_scene
    a = Renderer(ctx, sc, 48, 32, 6)
    a.set_resolution(48, 32)
    a.set_options(rr_start=3, seed=7)
    a.set_target_spp(5)
    a.start()
    a.wait()
    assert a.current_sample_count() == 5
    assert not a.is_running()
    b = Renderer(ctx, sc, 48, 32, 6)
    b.set_resolution(48, 32)
    b.set_options(rr_start=3, seed=7)
    b.render(5)
    assert np.array_equal(a.raw(), b.raw())
    st = a.current_stats()
    assert st["samples"] == 5 and st["total_rays"] > 0 and st["running_time"] > 0
    a.close()
    b.close()


def test_start_clears_and_pause_stops(cornell_scene):
    ctx, sc = cornell_scene
    r = Renderer(ctx, sc, 32, 32, 4)
    r.set_target_spp(0)              # unlimited, like the reference default
    r.start()
    with pytest.raises(CrtError, match="STATE"):
        r.wait()                      # would never return: rejected instead of hanging
    import time
    time.sleep(0.2)
    r.pause()
    n = r.current_sample_count()
    assert n >= 1
    time.sleep(0.05)
    assert r.current_sample_count() == n      # stopped
    r.wait()                                  # paused: returns at once
    r.set_target_spp(3)
    r.start()                                 # clears accumulation (renderer.cpp:159-163)
    r.wait()
    assert r.current_sample_count() == 3
    r.close()


def test_checkpoint_resume_bit_exact(cornell_scene):
    ctx, sc = cornell_scene
    full = Renderer(ctx, sc, 40, 24, 8)
    full.set_resolution(40, 24)
    full.set_options(rr_start=3, seed=11)
    full.render(6)
    part = Renderer(ctx, sc, 40, 24, 8)
    part.set_resolution(40, 24)
    part.set_options(rr_start=3, seed=11)
    part.render(2)
    blob = part.checkpoint()
    part.close()
    resumed = Renderer(ctx, sc, 40, 24, 8)
    resumed.set_resolution(40, 24)
    resumed.restore(blob)
    assert resumed.current_sample_count() == 2
    resumed.render(4)
    assert np.array_equal(resumed.raw(), full.raw())
    np.testing.assert_array_equal(resumed.current_progress(), full.current_progress())
    # a checkpoint of another resolution is rejected
    other = Renderer(ctx, sc, 8, 8, 8)
    with pytest.raises(CrtError):
        other.restore(blob)
    other.close()
    resumed.close()
    full.close()


def test_buffers_before_first_pass_are_flt_max(cornell_scene):
    ctx, sc = cornell_scene
    r = Renderer(ctx, sc, 8, 8, 2)
    big = np.float32(3.402823466e38)
    assert np.all(r.current_progress() == big)      # image.h:30-33
    assert np.all(r.current_normals() == big)
    assert np.all(r.raw() == 0.0)
    r.render(1)
    assert np.all(r.current_progress() <= 1.0)
    assert r.current_resolution() == (8, 8)
    r.set_resolution(16, 4)
    assert r.current_resolution() == (16, 4) and r.current_sample_count() == 0
    r.close()


def test_two_models_geom_ids():
    a = scenes.cornell(1)
    b = scenes.soup(500, seed=3)
    b.vertices = b.vertices + np.float32(5.0)          # far away from the box
    ctx = Context(0)
    sc = Scene(ctx)
    ma = sc.add_model(a)
    mb = sc.add_model(b)
    assert (ma, mb) == (0, 1)
    sc.set_camera(a.camera)
    sc.commit()
    o1, d1 = scenes.rays_from_camera(a, (16, 16))
    o2, d2 = scenes.diffuse_rays(4000, seed=5, lo=5.0, hi=6.0)
    got = sc.cast_rays(np.concatenate([o1, o2]), np.concatenate([d1, d2]))
    g, p = got["geom_id"], got["prim_id"]
    hit = p != MISS
    assert np.all(g[:256][hit[:256]] == 0) and np.all(p[:256][hit[:256]] < a.ntri)
    assert np.any(g[256:] == 1) and np.all(p[256:][g[256:] == 1] < b.ntri)
    sc.close()
    ctx.close()


def test_error_paths():
    ctx = Context(0)
    sc = Scene(ctx)
    bad = scenes.cornell(1)
    bad.vertices = bad.vertices.copy()
    bad.vertices[3, 1] = np.nan
    with pytest.raises(CrtError, match="INVALID_MESH"):
        sc.add_model(bad)
    bad = scenes.cornell(1)
    bad.material_indices = bad.material_indices.copy()
    bad.material_indices[0] = 99
    with pytest.raises(CrtError, match="INVALID_MESH"):
        sc.add_model(bad)
    bad = scenes.cornell(1)
    bad.vertex_indices = bad.vertex_indices.copy()
    bad.vertex_indices[5] = 10 ** 6
    with pytest.raises(CrtError, match="INVALID_MESH"):
        sc.add_model(bad)
    with pytest.raises(CrtError, match="STATE"):
        sc.cast_rays(np.zeros((1, 3)), np.ones((1, 3)))        # not committed
    sc.add_model(scenes.cornell(1))
    with pytest.raises(CrtError, match="INVALID_ARG"):
        sc.set_instances(0, np.eye(4).reshape(-1) * 2)          # bottom row (0, 0, 0, 2): not affine
    singular = np.eye(4)
    singular[1, 1] = 0.0
    with pytest.raises(CrtError, match="INVALID_ARG"):
        sc.set_instances(0, singular.reshape(-1))
    with pytest.raises(CrtError, match="INVALID_ARG"):
        sc.set_instances(0, np.full(16, np.nan))
    sc.close()
    ctx.close()


def test_destroy_refuses_live_dependents():
    """crt_context_destroy refuses a context with live scenes / renderers / caches, and
    crt_nrc_destroy / crt_scene_destroy refuse an object a renderer still uses;
    Context.close() closes what is left in dependency order."""
    from crender_amd import Nrc
    from crender_amd import _abi as A

    L = A.lib()
    STATE = -5
    ctx = Context(0)
    sc = scene_from_spec(ctx, scenes.cornell(1))
    r = Renderer(ctx, sc, 16, 16, 4)
    nrc = Nrc(ctx, max_batch=4096, seed=1)
    r.set_nrc(nrc, train_paths=4096)
    assert L.crt_context_destroy(ctx.handle) == STATE
    assert L.crt_nrc_destroy(nrc.handle) == STATE
    assert L.crt_scene_destroy(sc.handle) == STATE
    r.set_nrc(None)
    nrc.close()                       # detached: destroy succeeds
    r.render(1)
    ctx.close()                       # closes the renderer, then the scene
    assert r._h is None and sc.handle is None


@pytest.mark.parametrize("mode", [0, 1])
def test_multi_pass_wavefronts_equal_single_passes(mode):
    """A wavefront carrying several sample passes (the default for small images: up to 8)
    gives the same image, AOVs and counts as one pass per wavefront, for a pass count that
    is not a multiple of the wavefront's, synchronously and in the background."""
    spec = scenes.sunlit(3, mode=mode)
    res = (40, 30)
    out = {}
    for name, wf in (("one", 1), ("many", 0)):
        ctx = Context(0, wavefront_paths=wf)
        sc = scene_from_spec(ctx, spec)
        r = Renderer(ctx, sc, res[0], res[1], 5)
        r.set_resolution(*res)
        r.set_options(mode=mode, rr_start=2, seed=9)
        r.render(5)
        r.render(6)
        bg = Renderer(ctx, sc, res[0], res[1], 5)
        bg.set_resolution(*res)
        bg.set_options(mode=mode, rr_start=2, seed=9)
        bg.set_target_spp(11)
        bg.start()
        bg.wait()
        st = r.current_stats()
        out[name] = (r.raw(), r.current_normals(), r.current_albedos(), r.current_depths(), bg.raw(),
                     r.current_sample_count(), bg.current_sample_count(), st["extension_rays"], st["shadow_rays"])
        ctx.close()
    a, b = out["one"], out["many"]
    for x, y in zip(a[:5], b[:5]):
        assert np.array_equal(x, y)
    assert a[5:] == b[5:] and a[5] == 11 and a[6] == 11


def test_golden_rays():
    for name, spec in (("cornell_rays", scenes.cornell(1)), ("soup_rays", scenes.soup(20_000, seed=13))):
        g = np.load(GOLDEN / f"{name}.npz")
        ctx = Context(0)
        sc = scene_from_spec(ctx, spec)
        got = sc.cast_rays(g["org"], g["dir"])
        p = got["prim_id"].astype(np.int64)
        p[p == MISS] = -1
        mism = np.nonzero(p != g["prim"])[0]
        assert len(mism) <= max(1, int(1e-5 * len(p))), f"{name}: {len(mism)} mismatches"
        h = (p >= 0) & (p == g["prim"])
        assert np.max(np.abs(got["t"][h] - g["t"][h]) / np.maximum(g["t"][h], 1e-6)) < 1e-5
        sc.close()
        ctx.close()


def test_golden_cornell_image():
    g = np.load(GOLDEN / "cornell_compat.npz")
    spec = scenes.cornell(1)
    ctx = Context(0)
    sc = scene_from_spec(ctx, spec)
    r = Renderer(ctx, sc, 64, 64, spec.max_bounces)
    r.set_resolution(64, 64)
    r.set_options(mode=0, rr_start=spec.rr_start, seed=5)
    r.render(4)
    raw = r.raw().astype(np.float64)
    ref = g["raw"]
    close = np.isclose(raw, ref, rtol=1e-3, atol=1e-4).all(axis=-1).mean()
    assert close >= 0.99
    assert abs(raw.mean() - ref.mean()) <= 5e-3 * ref.mean()
    r.close()
    sc.close()
    ctx.close()


def test_cpp_host_example_runs():
    """The C++ host mirror (crender_amd/host/crt.hpp) driving the library end to end."""
    import subprocess

    exe = Path(__file__).resolve().parent.parent / "crender_amd" / "examples" / "render_box"
    assert exe.exists(), "build() compiles the C++ example"
    out = subprocess.run([str(exe), "8"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "samples 8," in out.stdout
    assert "cast_ray: distance 4.4000 material 2 normal 0.0 0.0 -1.0 point 0.00 1.00 1.00" in out.stdout
    mean = float(out.stdout.strip().split("mean display ")[-1])
    assert 0.05 < mean < 1.0
