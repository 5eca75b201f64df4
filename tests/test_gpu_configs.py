"""The batched and terrain workloads at their own sizes, against the CPU oracle.

  * batched rendering: four jobs of the 64-scene batch suite (soup and terrain, 18k-93k
    triangles, 128^2 and 256^2, 4 and 16 spp, depth 6) rendered together by
    crt_render_batch; each job's sum is compared with the oracle's render of that scene
    (FP32 BVH) at the image-parity bar of tests/_parity.py;
  * terrain with objects at full size (2047^2 * 2 terrain triangles + 20k icospheres, 14.8M
    triangles, 1920x1080, pbr, sun + sky, depth 8): a band of 16 rows at 4 spp against the
    oracle's render of the same rows (same Philox streams), 20k sampled closest-hit queries
    against the oracle's BVH, and properties of a 64-spp render (finite, non-negative, ray
    counts within their bounds).
"""
import numpy as np
import pytest

from _parity import assert_parity
from crender_amd import Context, Renderer, render_batch, scene_from_spec, scenes

pytestmark = pytest.mark.gpu

BATCH_JOBS = (3, 12, 31, 36)    # indices into scenes.batch_suite_params(64, 7)
MISS = 0xFFFFFFFF


def test_batch_suite_jobs_match_oracle(oracle):
    params = [p for p in scenes.batch_suite_params(64, 7) if p["index"] in BATCH_JOBS]
    assert all(p["tris"] <= 100_000 and p["res"] <= 256 for p in params)
    specs = [scenes.batch_job_scene(p) for p in params]
    ctx = Context(0)
    scs = [scene_from_spec(ctx, s) for s in specs]
    jobs = [dict(scene=sc, width=p["res"], height=p["res"], spp=p["spp"], max_bounces=6, mode=s.mode, rr_start=0,
                 seed=p["index"]) for sc, p, s in zip(scs, params, specs)]
    print("batch jobs", [(p["kind"], p["tris"], p["res"], p["spp"]) for p in params], flush=True)
    outs, _, st = render_batch(ctx, jobs)
    ext = 0
    for out, p, s in zip(outs, params, specs):
        kw = dict(res=(p["res"], p["res"]), spp=p["spp"], max_bounces=6, rr_start=0, mode=s.mode, seed=p["index"],
                  aovs=False)
        bvh = oracle.Bvh(s.corners())
        ref = oracle.render(s, trav=oracle.BVH32, bvh=bvh, **kw)
        print(f"job {p['index']}", flush=True)
        assert_parity(out, ref["raw"], lambda: oracle.render(s, trav=oracle.BVH64, bvh=bvh, **kw)["raw"])
        ext += ref["stats"]["ext_rays"]
    assert abs(st["extension_rays"] - ext) <= 1e-3 * ext
    for sc in scs:
        sc.close()
    ctx.close()


@pytest.fixture(scope="module")
def terrain_full(oracle):
    spec = scenes.terrain()            # 2048^2 grid, 20k icospheres, 1920x1080
    assert spec.ntri == 14_780_418
    bvh = oracle.Bvh(spec.corners())
    return spec, bvh


def test_terrain_full_size_band_matches_oracle(oracle, terrain_full):
    spec, bvh = terrain_full
    W, H = spec.res
    rows, spp, seed = (532, 548), 4, 5
    print("terrain full size, band of rows", flush=True)
    ctx = Context(0)
    sc = scene_from_spec(ctx, spec)
    r = Renderer(ctx, sc, W, H, spec.max_bounces)
    r.set_resolution(W, H)
    r.set_options(mode=spec.mode, rr_start=spec.rr_start, seed=seed)
    r.render(spp)
    gpu = r.raw()
    r.close()
    kw = dict(res=(W, H), spp=spp, seed=seed, rows=rows, aovs=False)
    ref = oracle.render(spec, trav=oracle.BVH32, bvh=bvh, **kw)["raw"]
    # rows y of the render land at H - 1 - y of the (flipped) sum
    band = slice(H - rows[1], H - rows[0])
    assert np.all(ref[:H - rows[1]] == 0) and np.all(ref[H - rows[0]:] == 0)
    assert_parity(gpu[band], ref[band], lambda: oracle.render(spec, trav=oracle.BVH64, bvh=bvh, **kw)["raw"][band])
    sc.close()
    ctx.close()


def test_terrain_full_size_hits_and_properties(oracle, terrain_full):
    spec, bvh = terrain_full
    W, H = spec.res
    ctx = Context(0)
    sc = scene_from_spec(ctx, spec)
    # closest hits: camera rays at 10k random pixels + 10k rays from points above the terrain
    o1, d1 = scenes.rays_from_camera(spec, (W, H), seed=3)
    pick = np.random.default_rng(4).choice(len(o1), 10_000, replace=False)
    o2, d2 = scenes.diffuse_rays(10_000, seed=14, lo=0.05, hi=0.95)
    o2[:, 1] = 0.25 + 0.1 * o2[:, 1]                     # above the highest terrain point
    o = np.concatenate([o1[pick], o2])
    d = np.concatenate([d1[pick], d2])
    print("terrain full size, 20k closest-hit queries", flush=True)
    got = sc.cast_rays(o, d)
    gp = got["prim_id"].astype(np.int64)
    gp[gp == MISS] = -1
    pb, tb, _, _ = oracle.intersect(spec.corners(), o, d, oracle.BVH32, bvh=bvh)
    assert int(np.sum(gp != pb)) <= max(1, int(1e-5 * len(gp)))
    hit = (pb >= 0) & (gp == pb)
    assert hit.mean() > 0.3      # camera rays above the horizon and upward rays miss
    assert np.max(np.abs(got["t"][hit] - tb[hit]) / np.maximum(tb[hit], 1e-6)) <= 1e-5
    # a 64-spp render at full size
    print("terrain full size, 64 spp", flush=True)
    r = Renderer(ctx, sc, W, H, spec.max_bounces)
    r.set_resolution(W, H)
    r.set_options(mode=spec.mode, rr_start=spec.rr_start, seed=1)
    r.render(spec.spp)
    img = r.raw() / spec.spp
    st = r.current_stats()
    r.close()
    assert np.isfinite(img).all() and img.min() >= 0.0
    assert img.mean() > 0.05
    paths = W * H * spec.spp
    assert st["samples"] == spec.spp
    assert paths <= st["extension_rays"] <= paths * spec.max_bounces
    assert 0 < st["shadow_rays"] <= st["extension_rays"]
    assert st["extension_rays"] <= st["total_rays"] <= paths * (spec.max_bounces + 1)
    sc.close()
    ctx.close()
