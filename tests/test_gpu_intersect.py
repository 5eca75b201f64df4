"""Closest-hit parity of the HIP traversal (crt_intersect) against the CPU oracle.

Bar: primID bit-exact except near ties; t within 1e-5 relative.  Against the oracle's
FP32 watertight test (same arithmetic) the agreement is exact; against FP64 brute force
every disagreement must be a near tie (the two triangles' FP64 distances within 1e-5
relative, or the FP64 hit grazing an edge) and at most 1e-5 of the rays.
"""
import numpy as np
import pytest

from crender_amd import Context, scene_from_spec, scenes

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def _prims(got):
    p = got["prim_id"].astype(np.int64)
    p[p == MISS] = -1
    return p


def _classify_fp64(tri, o, d, gp, p64, t64):
    """Returns the number of disagreements that are NOT near ties."""
    import pyoracle as O

    bad = 0
    for i in np.nonzero(gp != p64)[0]:
        if gp[i] < 0 or p64[i] < 0:
            # a hit missed by one side must be a grazing hit in FP64
            k = p64[i] if p64[i] >= 0 else gp[i]
            pt, tt, uu, vv = O.intersect(tri[k:k + 1], o[i:i + 1], d[i:i + 1], O.BRUTE64)
            if pt[0] < 0 or min(uu[0], vv[0], 1 - uu[0] - vv[0]) > 1e-5:
                bad += 1
            continue
        # distance of the GPU's triangle in FP64
        pt, tt, _, _ = O.intersect(tri[gp[i]:gp[i] + 1], o[i:i + 1], d[i:i + 1], O.BRUTE64)
        if pt[0] < 0 or abs(tt[0] - t64[i]) > 1e-5 * max(t64[i], 1e-3):
            bad += 1
    return bad


@pytest.mark.parametrize("width", [4, 8])
def test_cornell_all_rays_vs_fp64(oracle, width):
    spec = scenes.cornell(1)
    ctx = Context(0, bvh_width=width)
    sc = scene_from_spec(ctx, spec)
    o, d = scenes.rays_from_camera(spec, (128, 128))
    got = sc.cast_rays(o, d, normals=True)
    tri = spec.corners()
    p64, t64, u64, v64 = oracle.intersect(tri, o, d, oracle.BRUTE64)
    p32, t32, u32, v32 = oracle.intersect(tri, o, d, oracle.BRUTE32)
    gp = _prims(got)
    # FP32 watertight: identical arithmetic -> bit-exact
    assert np.array_equal(gp, p32)
    hit = p32 >= 0
    assert np.array_equal(got["t"][hit], t32[hit].astype(np.float32))
    assert np.array_equal(got["u"][hit], u32[hit].astype(np.float32))
    assert np.array_equal(got["v"][hit], v32[hit].astype(np.float32))
    assert np.all(np.isinf(got["t"][~hit]))
    # FP64 truth
    assert _classify_fp64(tri, o, d, gp, p64, t64) == 0
    h = (gp >= 0) & (p64 >= 0)
    assert np.max(np.abs(got["t"][h] - t64[h]) / np.maximum(t64[h], 1e-6)) < 1e-5
    # Ng = (v1-v0)x(v2-v0)
    c = tri.reshape(-1, 3, 3).astype(np.float64)
    ng = np.cross(c[:, 1] - c[:, 0], c[:, 2] - c[:, 0])[gp[h]]
    np.testing.assert_allclose(np.stack([got["ng_x"], got["ng_y"], got["ng_z"]], 1)[h], ng, rtol=1e-5, atol=1e-6)
    assert np.all(got["geom_id"][h] == 0) and np.all(got["geom_id"][~h] == MISS)
    sc.close()
    ctx.close()


@pytest.mark.parametrize("width", [4, 8])
def test_soup_100k_vs_oracle(oracle, width):
    spec = scenes.soup(100_000, seed=5)
    tri = spec.corners()
    ctx = Context(0, bvh_width=width)
    sc = scene_from_spec(ctx, spec)
    info = sc.bvh_info()
    assert info["width"] == width and info["node_count"] > 0
    o1, d1 = scenes.rays_from_camera(spec, (320, 180))
    o2, d2 = scenes.diffuse_rays(60_000, seed=11)
    o = np.concatenate([o1, o2])
    d = np.concatenate([d1, d2])
    got = sc.cast_rays(o, d)
    gp = _prims(got)
    bvh = oracle.Bvh(tri)
    pb, tb, ub, vb = oracle.intersect(tri, o, d, oracle.BVH32, bvh=bvh)
    assert np.array_equal(gp, pb), f"{np.sum(gp != pb)} of {len(gp)} differ from the FP32 oracle"
    hit = pb >= 0
    assert np.array_equal(got["t"][hit], tb[hit].astype(np.float32))
    # FP64 brute force on a sample of 2000 rays
    sel = np.linspace(0, len(o) - 1, 2000).astype(np.int64)
    p64, t64, _, _ = oracle.intersect(tri, o[sel], d[sel], oracle.BRUTE64)
    bad = _classify_fp64(tri, o[sel], d[sel], gp[sel], p64, t64)
    assert bad == 0
    assert np.mean(gp[sel] != p64) <= max(1e-5, 1.0 / len(sel))
    sc.close()
    ctx.close()


@pytest.fixture(scope="module")
def soup_1m(oracle):
    """The headline workload's 1M-triangle soup, primary rays of its 1920x1080 camera (every
    16th pixel) plus diffuse rays from inside the cloud, and the oracle's answers."""
    spec = scenes.soup(1_000_000, seed=2)
    tri = spec.corners()
    o1, d1 = scenes.rays_from_camera(spec, (1920, 1080))
    o1, d1 = o1[::16], d1[::16]
    o2, d2 = scenes.diffuse_rays(40_000, seed=12)
    o = np.concatenate([o1, o2])
    d = np.concatenate([d1, d2])
    bvh = oracle.Bvh(tri)
    pb, tb, _, _ = oracle.intersect(tri, o, d, oracle.BVH32, bvh=bvh)
    sel = np.linspace(0, len(o) - 1, 400).astype(np.int64)
    p64, t64, _, _ = oracle.intersect(tri, o[sel], d[sel], oracle.BRUTE64)
    return spec, tri, o, d, pb, tb, sel, p64, t64


@pytest.mark.parametrize("stack", [8, 16, 32])
def test_soup_1m_headline_scale_vs_oracle(soup_1m, stack):
    """Hit-ID agreement at the headline size for every LDS stack depth the traversal is built
    with (8 and 16 entries spill to the global stack far more often on this deep BVH; the
    renderers that train the radiance cache trace with 16)."""
    spec, tri, o, d, pb, tb, sel, p64, t64 = soup_1m
    print(f"soup 1M, trace_stack {stack}", flush=True)
    ctx = Context(0, trace_stack=stack)
    sc = scene_from_spec(ctx, spec)
    got = sc.cast_rays(o, d)
    gp = _prims(got)
    mism = int(np.sum(gp != pb))
    assert mism <= max(1, int(1e-5 * len(gp))), f"{mism} of {len(gp)} differ from the FP32 oracle"
    hit = (pb >= 0) & (gp == pb)
    assert np.max(np.abs(got["t"][hit] - tb[hit]) / np.maximum(tb[hit], 1e-6)) <= 1e-5
    assert _classify_fp64(tri, o[sel], d[sel], gp[sel], p64, t64) == 0
    sc.close()
    ctx.close()


@pytest.mark.parametrize("opts", [dict(trace_refill=8, trace_park_min=2), dict(trace_leaf_rounds=1),
                                  dict(trace_leaf_rounds=100, trace_park_min=48),
                                  dict(bvh_prim_cost=0.2, max_leaf=8), dict(bvh_prim_cost=5.0)],
                         ids=["refill8-park2", "leaf1", "leafall-park48", "primcost0.2-leaf8", "primcost5"])
def test_soup_1m_traversal_options_vs_oracle(soup_1m, opts):
    """The traversal's scheduling thresholds and the collapse's cost model change the order of
    work and the tree, never the answers: the same bar as the default settings."""
    spec, tri, o, d, pb, tb, sel, p64, t64 = soup_1m
    print(f"soup 1M, {opts}", flush=True)
    ctx = Context(0, **opts)
    sc = scene_from_spec(ctx, spec)
    got = sc.cast_rays(o, d)
    gp = _prims(got)
    mism = int(np.sum(gp != pb))
    assert mism <= max(1, int(1e-5 * len(gp))), f"{mism} of {len(gp)} differ from the FP32 oracle"
    hit = (pb >= 0) & (gp == pb)
    assert np.max(np.abs(got["t"][hit] - tb[hit]) / np.maximum(tb[hit], 1e-6)) <= 1e-5
    sc.close()
    ctx.close()


def test_edge_cases(oracle):
    spec = scenes.cornell(2)
    ctx = Context(0)
    sc = scene_from_spec(ctx, spec)
    # empty batch
    out = sc.cast_rays(np.zeros((0, 3)), np.zeros((0, 3)))
    assert out["t"].shape == (0,)
    # tnear > tfar, NaN direction, zero direction, axis-parallel rays through walls
    o = np.array([[0, 1, 0], [0, 1, 0], [0, 1, 0], [0, 1, 0], [0, 1, 0], [0.0, 1.0, 0.0]], np.float32)
    d = np.array([[0, 0, 1], [np.nan, 0, 1], [0, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0]], np.float32)
    tn = np.array([5, 1e-5, 1e-5, 1e-5, 1e-5, 1e-5], np.float32)
    tf = np.array([1, np.inf, np.inf, np.inf, np.inf, 0.5], np.float32)
    got = sc.cast_rays(o, d, tnear=tn, tfar=tf)
    gp = _prims(got)
    assert gp[0] == -1 and gp[1] == -1 and gp[2] == -1
    p32, t32, _, _ = oracle.intersect(spec.corners(), o[3:], d[3:], oracle.BRUTE32, tnear=tn[3:], tfar=tf[3:])
    assert np.array_equal(gp[3:], p32)
    assert abs(got["t"][3] - 1.0) < 1e-6 and abs(got["t"][4] - 1.0) < 1e-6   # right wall, floor
    assert gp[5] == -1   # ceiling at t = 1 beyond tfar = 0.5
    # occlusion agrees with closest hit
    o2, d2 = scenes.rays_from_camera(spec, (64, 64))
    occ = sc.occluded(o2, d2, tfar=3.0)
    ch = sc.cast_rays(o2, d2, tfar=3.0)
    assert np.array_equal(occ, ch["prim_id"] != MISS)
    sc.close()
    ctx.close()


def test_scene_with_no_triangles_misses_everything():
    spec = scenes.cornell(1)
    spec.vertices = spec.vertices[:0]
    spec.vertex_indices = spec.vertex_indices[:0]
    spec.material_indices = spec.material_indices[:0]
    spec.texture_indices = spec.texture_indices[:0]
    ctx = Context(0)
    sc = scene_from_spec(ctx, spec)
    got = sc.cast_rays(np.zeros((5, 3)), np.tile([0, 0, 1.0], (5, 1)))
    assert np.all(got["prim_id"] == MISS) and np.all(np.isinf(got["t"]))
    sc.close()
    ctx.close()
