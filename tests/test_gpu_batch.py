"""Batched rendering (crt_render_batch) against the single-scene renderer.

A job's running sum must equal, bit for bit, what a renderer with the same scene,
resolution, bounces and options accumulates over the same passes: the batch uses the
same Philox streams (seed, pixel, sample, bounce), the same shading code and sums each
pixel's passes in pass order.  Covered: several scenes of different size and material mix
(sun + shadow rays, compat and pbr), resolutions, spp and depths; wavefront sizes that
force many replays and split one image over several replays; graph replay and direct
launches.
"""
import numpy as np
import pytest

from crender_amd import Context, Renderer, render_batch, scene_from_spec, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    ctx = Context(0)
    specs = [scenes.cornell(1), scenes.soup(20_000, seed=5), scenes.sunlit(4, mode=1), scenes.soup(3_000, seed=6),
             scenes.sunlit(5, mode=0)]
    scs = [scene_from_spec(ctx, s) for s in specs]
    jobs = [
        dict(scene=scs[0], width=64, height=48, spp=3, max_bounces=8, mode=0, rr_start=3, seed=11),
        dict(scene=scs[1], width=80, height=40, spp=1, max_bounces=4, mode=0, rr_start=0, seed=12),
        dict(scene=scs[2], width=96, height=64, spp=4, max_bounces=6, mode=1, rr_start=3, seed=13),
        dict(scene=scs[3], width=33, height=17, spp=5, max_bounces=3, mode=1, rr_start=0, seed=14,
             sample_offset=7),
        dict(scene=scs[4], width=50, height=70, spp=2, max_bounces=5, mode=0, rr_start=2, seed=15),
    ]
    singles, ext = [], 0
    for j in jobs:
        r = Renderer(ctx, j["scene"], j["width"], j["height"], j["max_bounces"])
        r.set_resolution(j["width"], j["height"])
        r.set_options(mode=j["mode"], rr_start=j["rr_start"], seed=j["seed"], sample_offset=j.get("sample_offset", 0))
        r.render(j["spp"])
        singles.append(r.raw())
        ext += r.current_stats()["extension_rays"]
        r.close()
    yield ctx, jobs, singles, ext
    for s in scs:
        s.close()
    ctx.close()


@pytest.mark.parametrize("max_paths,use_graph", [(0, True), (0, False), (4096, True), (2560, False), (1024, True)])
def test_batch_equals_single_renderer(setup, max_paths, use_graph):
    ctx, jobs, singles, ext = setup
    print(f"batch max_paths={max_paths} graph={use_graph}", flush=True)
    outs, ms, st = render_batch(ctx, jobs, max_paths=max_paths, use_graph=use_graph)
    for i, (a, b) in enumerate(zip(outs, singles)):
        assert a.shape == b.shape
        assert np.array_equal(a, b), (i, np.abs(a - b).max())
    assert st["extension_rays"] == ext
    assert st["used_graph"] == (1 if use_graph else 0)
    if max_paths:
        assert st["replays"] > 1
    # jobs are packed in order, so they finish in order
    assert all(ms[i] <= ms[i + 1] + 1e-6 for i in range(len(ms) - 1)), ms
    assert st["rays_per_second"] > 0


def test_batch_graph_replays_repeat(setup):
    ctx, jobs, singles, _ = setup
    for _ in range(2):
        outs, _, st = render_batch(ctx, jobs, max_paths=3000, use_graph=True)
        assert st["replays"] > 3
        for a, b in zip(outs, singles):
            assert np.array_equal(a, b)


def test_batch_rejects_bad_jobs(setup):
    ctx, jobs, _, _ = setup
    from crender_amd import CrtError
    bad = dict(jobs[0])
    bad["spp"] = 0
    with pytest.raises(CrtError):
        render_batch(ctx, [bad])
    bad = dict(jobs[0])
    bad["max_bounces"] = 0
    with pytest.raises(CrtError):
        render_batch(ctx, [bad])


def test_autotune_then_batch_still_exact():
    from crender_amd import autotune
    ctx = Context(0)
    try:
        specs = [scenes.soup(10_000, seed=7), scenes.sunlit(6, mode=1)]
        scs = [scene_from_spec(ctx, s) for s in specs]
        jobs = [dict(scene=scs[0], width=64, height=64, spp=2, max_bounces=4, mode=0, seed=3),
                dict(scene=scs[1], width=48, height=32, spp=3, max_bounces=5, mode=1, rr_start=2, seed=4)]
        res = autotune(ctx, jobs, tune_width=True)
        print("autotune", res, flush=True)
        assert res["bvh_width"] in (4, 8) and res["trace_stack"] in (8, 16, 32)
        assert 2 <= res["trace_blocks_per_cu"] <= 8 and res["max_paths"] >= (1 << 20)
        assert res["trials"] >= 8 and res["rays_per_second"] > 0
        outs, _, st = render_batch(ctx, jobs)
        assert st["bvh_width"] == res["bvh_width"]
        for j, o in zip(jobs, outs):
            r = Renderer(ctx, j["scene"], j["width"], j["height"], j["max_bounces"])
            r.set_resolution(j["width"], j["height"])
            r.set_options(mode=j["mode"], rr_start=j.get("rr_start", 0), seed=j["seed"])
            r.render(j["spp"])
            assert np.array_equal(r.raw(), o)
            r.close()
        for s in scs:
            s.close()
    finally:
        ctx.close()


def test_autotune_beside_a_background_renderer():
    """The autotuner re-sizes the context's traversal grid (and its spill area) while a
    renderer of the same context renders in the background: the passes and the tuner's
    batches take turns on the context, and the background image is the synchronous one."""
    from crender_amd import autotune
    ctx = Context(0)
    try:
        soup = scene_from_spec(ctx, scenes.soup(20_000, seed=9))
        box = scene_from_spec(ctx, scenes.cornell(1))
        bg = Renderer(ctx, box, 96, 96, 6)
        bg.set_resolution(96, 96)
        bg.set_options(mode=0, rr_start=3, seed=12)
        bg.set_target_spp(40)
        bg.start()
        jobs = [dict(scene=soup, width=64, height=64, spp=2, max_bounces=4, mode=0, seed=3)]
        res = autotune(ctx, jobs)
        assert res["trials"] >= 4
        bg.wait()
        assert bg.current_sample_count() == 40
        ref = Renderer(ctx, box, 96, 96, 6)
        ref.set_resolution(96, 96)
        ref.set_options(mode=0, rr_start=3, seed=12)
        ref.render(40)
        assert np.array_equal(bg.raw(), ref.raw())
        bg.close()
        ref.close()
        soup.close()
        box.close()
    finally:
        ctx.close()
