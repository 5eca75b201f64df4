"""crt_cast_rays (scene::cast_ray returning the intersection record, scene.cpp:79-98,
model.cpp:27-47) on the GPU:

  * the closed forms of the two-quad scene (distance, point, normal, uv, material, triangle,
    misses, tnear) that also pin the oracle;
  * field by field, bit for bit, against the oracle's records (FP32, brute force) on the
    Cornell box and on the textured cut-out scene (uv interpolated over negative and > 1
    coordinates), camera rays of every pixel plus rays from random points;
  * device memory with the reference's default tnear / tfar (NULL pointers) and only some
    fields requested: equal to the host-memory call;
  * a two-level scene: consistent with crt_intersect (same hit, distance = t |d|, normal =
    the normalised object-space Ng, point = o + d t).
"""
import numpy as np
import pytest

from _devmem import DeviceArrays
from _records import check_miss, check_oblique, check_quad_a, check_quad_b, quad_a_rays, quad_b_rays, quads_spec
from crender_amd import Context, Scene, scene_from_spec, scenes

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def _as_rec(g):
    ids = lambda a: np.where(a == MISS, -1, a.astype(np.int64))   # noqa: E731
    return {"distance": g["distance"], "prim": ids(g["prim_id"]), "material": ids(g["material"]),
            "uv": np.column_stack([g["u"], g["v"]]), "normal": np.column_stack([g["n_x"], g["n_y"], g["n_z"]]),
            "point": np.column_stack([g["p_x"], g["p_y"], g["p_z"]])}


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def test_closed_forms(ctx):
    sc = scene_from_spec(ctx, quads_spec())
    o, d, xy = quad_a_rays()
    print("records: quad A", len(o), flush=True)
    check_quad_a(_as_rec(sc.cast_ray_records(o, d)), xy)
    o, d = quad_b_rays()
    check_quad_b(_as_rec(sc.cast_ray_records(o, d)), o)
    ab = np.random.default_rng(2).uniform(-0.9, 0.9, (300, 2)).astype(np.float32)
    ab = ab[np.abs(ab[:, 0] - ab[:, 1]) > 0.01]
    d = np.column_stack([ab, np.full(len(ab), 2.0, np.float32)])
    check_oblique(_as_rec(sc.cast_ray_records(np.zeros_like(d), d)), ab)
    check_miss(_as_rec(sc.cast_ray_records(np.zeros((3, 3), np.float32),
                                           np.array([(0, 0, -1), (-1, 0, 0), (0, 1, 0)], np.float32))))
    for below, hits in ((0.0, False), (5e-6, False), (1e-3, True)):
        r = _as_rec(sc.cast_ray_records(np.array([(0.3, -0.2, 2.0 - below)], np.float32),
                                        np.array([(0, 0, 1)], np.float32)))
        assert (r["prim"][0] == 0) if hits else np.isinf(r["distance"][0])
    sc.close()


@pytest.mark.parametrize("name", ["cornell", "cutout"])
def test_records_match_oracle(ctx, oracle, name):
    spec = getattr(scenes, name)()
    sc = scene_from_spec(ctx, spec)
    o1, d1 = scenes.rays_from_camera(spec, (96, 64), seed=3)
    lo, hi = spec.vertices.min(axis=0), spec.vertices.max(axis=0)
    o2, d2 = scenes.diffuse_rays(20_000, seed=8)
    o2 = (lo + (hi - lo) * (0.05 + 0.9 * o2)).astype(np.float32)
    o = np.concatenate([o1, o2]).astype(np.float32)
    d = np.concatenate([d1, d2]).astype(np.float32)
    print(f"records vs oracle: {name}, {len(o)} rays", flush=True)
    g = _as_rec(sc.cast_ray_records(o, d))
    ref = oracle.cast_records(spec, o, d, trav=oracle.BRUTE32)
    r = {"distance": ref["distance"], "prim": ref["prim_id"], "material": ref["material"], "uv": ref["uv"],
         "normal": ref["normal"], "point": ref["point"]}
    hit = r["prim"] >= 0
    assert 0.2 < hit.mean()
    np.testing.assert_array_equal(g["prim"], r["prim"])
    for k in ("distance", "material", "uv", "normal", "point"):
        np.testing.assert_array_equal(g[k], r[k], err_msg=k)
    if name == "cutout":
        assert np.any(g["uv"][hit] < 0) and np.any(g["uv"][hit] > 1)   # wrap-around coordinates were hit
    sc.close()


def test_device_memory_and_default_limits(ctx):
    spec = scenes.cornell()
    sc = scene_from_spec(ctx, spec)
    o, d = scenes.rays_from_camera(spec, (64, 64), seed=4)
    o, d = o.astype(np.float32), d.astype(np.float32)
    n = len(o)
    host = sc.cast_ray_records(o, d)
    dev = DeviceArrays()
    try:
        keys = ("org_x", "org_y", "org_z", "dir_x", "dir_y", "dir_z")
        rays = {k: dev.put(c) for k, c in zip(keys, [o[:, 0], o[:, 1], o[:, 2], d[:, 0], d[:, 1], d[:, 2]])}
        want = {"distance": np.float32, "material": np.uint32, "p_x": np.float32, "n_z": np.float32}
        out = {k: dev.put(np.full(n, 7, t)) for k, t in want.items()}
        print("records: device memory", n, flush=True)
        sc.cast_ray_records_device(rays, out, n)
        for k, t in want.items():
            np.testing.assert_array_equal(dev.get(out[k], n, t), host[k], err_msg=k)
    finally:
        dev.free()
    sc.close()


def test_instanced_records_consistent_with_hits(ctx):
    a = scenes.cornell()
    sc = Scene(ctx)
    ida = sc.add_model(a)
    T = np.eye(4, dtype=np.float32)
    T[:3, :3] = np.array([[0, -0.8, 0], [0.8, 0, 0], [0, 0, 0.8]], np.float32)   # 90 deg about z, scale 0.8
    T[:3, 3] = (0.4, 0.1, 0.13)   # no wall of this copy is coplanar with one of the other
    sc.set_instances(ida, [np.eye(4, dtype=np.float32).T.reshape(-1), T.T.reshape(-1)])
    sc.set_skybox(scenes.constant_sky(1.0))
    sc.set_camera(scenes.camera_dict((0.3, 0.4, -3.0), fov=60.0))
    sc.commit()
    lo, hi = a.vertices.min(axis=0), a.vertices.max(axis=0)
    o, d = scenes.diffuse_rays(20_000, seed=12)
    o = (lo + (hi - lo) * (0.05 + 0.9 * o)).astype(np.float32)
    d = (d * np.random.default_rng(5).uniform(0.5, 2.0, (len(d), 1))).astype(np.float32)   # not unit length
    print("records: instanced scene", len(o), flush=True)
    rec = sc.cast_ray_records(o, d)
    hits = sc.cast_rays(o, d, normals=True)
    hit = hits["prim_id"] != MISS
    assert hit.mean() > 0.5 and np.bincount(hits["inst_id"][hit]).min() > 1000   # both instances hit
    # The Cornell model's inner boxes stand on its floor: a ray starting inside a box and
    # pointing down meets two coplanar triangles at the same distance, and which one wins the
    # tie depends on the last bit of t, so the triangle may differ there but not the distance.
    rhit = rec["prim_id"] != MISS
    assert np.array_equal(rhit, hit)
    dn = np.linalg.norm(d.astype(np.float64), axis=1)
    # |P - o| with P = o + d t rounds to the coordinates' ulp (1.2e-7 near 1) at small t
    np.testing.assert_allclose(rec["distance"][hit], hits["t"][hit] * dn[hit], rtol=2e-5, atol=5e-7)
    same = (rec["prim_id"] == hits["prim_id"]) & (rec["geom_id"] == hits["geom_id"])
    assert same[hit].mean() > 0.995
    both = hit & same
    ng = np.column_stack([hits["ng_x"], hits["ng_y"], hits["ng_z"]]).astype(np.float64)[both]
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    nr = np.column_stack([rec["n_x"], rec["n_y"], rec["n_z"]])[both]
    np.testing.assert_allclose(nr, ng, atol=1e-5)
    p = np.column_stack([rec["p_x"], rec["p_y"], rec["p_z"]])[both]
    np.testing.assert_allclose(p, o[both] + d[both] * hits["t"][both, None], atol=2e-5)
    assert np.all(rec["material"][both] == a.material_indices[rec["prim_id"][both]])
    assert np.all(np.isinf(rec["distance"][~hit]))
    sc.close()
