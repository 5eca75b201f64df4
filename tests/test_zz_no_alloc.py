"""crt_intersect_noalloc (the allocation-free query) against crt_intersect on device memory.

Two calls in a row on the same scene, the second with fewer rays (the scratch keeps the
first call's contents), each compared field by field, bit for bit, with crt_intersect's
results for the same device rays; plus the size limit, and a two-level (instanced) scene.
Device buffers come from the HIP runtime through ctypes (the test allocates; the entry
point does not).
"""
import numpy as np
import pytest

from _devmem import DeviceArrays
from crender_amd import Context, CrtError, scene_from_spec, scenes

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 20

HIT_TYPES = dict(t=np.float32, u=np.float32, v=np.float32, prim_id=np.int32, geom_id=np.int32, ng_x=np.float32,
                 ng_y=np.float32, ng_z=np.float32, inst_id=np.int32)


def _device_rays(dev, o, d, tnear=1e-5, tfar=np.inf):
    n = o.shape[0]
    cols = [o[:, 0], o[:, 1], o[:, 2], d[:, 0], d[:, 1], d[:, 2], np.full(n, tnear), np.full(n, tfar)]
    keys = ("org_x", "org_y", "org_z", "dir_x", "dir_y", "dir_z", "tnear", "tfar")
    return {k: dev.put(np.asarray(c, np.float32)) for k, c in zip(keys, cols)}


def _device_hits(dev, n, fill):
    return {k: dev.put(np.full(n, fill, dt)) for k, dt in HIT_TYPES.items()}


def _both(dev, sc, o, d, fill):
    n = o.shape[0]
    rays = _device_rays(dev, o, d)
    ref_p = _device_hits(dev, n, fill)
    sc.cast_rays_device(rays, ref_p, n)
    got_p = _device_hits(dev, n, fill + 1)
    sc.cast_rays_device(rays, got_p, n, noalloc=True)
    ref = {k: dev.get(ref_p[k], n, dt) for k, dt in HIT_TYPES.items()}
    got = {k: dev.get(got_p[k], n, dt) for k, dt in HIT_TYPES.items()}
    return ref, got


def test_noalloc_equals_intersect_twice():
    dev = DeviceArrays()
    print("noalloc: soup 200k, 300k then 40k rays", flush=True)
    spec = scenes.soup(200_000, seed=8)
    ctx = Context(0)
    sc = scene_from_spec(ctx, spec)
    for n, seed in ((300_000, 1), (40_000, 2)):
        o1, d1 = scenes.rays_from_camera(spec, (640, 480), seed=seed)
        o2, d2 = scenes.diffuse_rays(n, seed=seed + 10)
        perm = np.random.default_rng(seed).permutation(len(o1) + n)[:n]   # camera and diffuse rays mixed
        o = np.concatenate([o1, o2])[perm]
        d = np.concatenate([d1, d2])[perm]
        ref, got = _both(dev, sc, o, d, fill=7)
        for k in ref:
            np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{k} (n={n})")
        assert (ref["prim_id"] != -1).mean() > 0.2
    # beyond the scratch: a clear error, nothing traced
    rays = _device_rays(dev, np.zeros((1, 3), np.float32), np.ones((1, 3), np.float32))
    hp = _device_hits(dev, 1, 0)
    with pytest.raises(CrtError, match="at most"):
        sc.cast_rays_device(rays, hp, MAX_RAYS + 1, noalloc=True)
    dev.free()
    sc.close()
    ctx.close()


def test_noalloc_instanced_scene():
    dev = DeviceArrays()
    print("noalloc: instanced cornell", flush=True)
    spec = scenes.cornell(2)
    ctx = Context(0)
    sc = scene_from_spec(ctx, spec, commit=False)
    t0 = np.eye(4, dtype=np.float32).reshape(16)
    t1 = np.eye(4, dtype=np.float32)
    t1[0, 3] = 3.0                     # a second copy 3 units along +x (row-major -> column-major below)
    sc.set_instances(0, np.stack([t0, t1.T.reshape(16)]))
    sc.commit()
    o, d = scenes.rays_from_camera(spec, (128, 96), seed=3)
    ref, got = _both(dev, sc, o, d, fill=5)
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    dev.free()
    sc.close()
    ctx.close()
