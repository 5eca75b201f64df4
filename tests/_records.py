"""Closed-form intersection records on two axis-aligned quads, shared by the oracle's pin
(test_oracle_records.py) and the GPU's (test_gpu_records.py).  Record dicts hold distance,
prim, material, uv (n, 2), normal and point (n, 3); misses carry -1 IDs."""
import numpy as np

from crender_amd import scenes
from crender_amd.scenes import MaterialSpec


def quads_spec():
    spec = scenes.furnace()
    a = [(-1, -1, 2), (1, -1, 2), (1, 1, 2), (-1, 1, 2)]
    b = [(5, -1, -1), (5, 1, -1), (5, 1, 9), (5, -1, 9)]
    spec.vertices = np.array(a + b, np.float32)
    spec.vertex_indices = np.array([0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7], np.uint32)
    spec.material_indices = np.array([0, 0, 1, 1], np.uint32)
    spec.materials = [MaterialSpec(colour=(0.5, 0.5, 0.5, 1.0)), MaterialSpec(colour=(0.2, 0.7, 0.2, 1.0))]
    spec.texture_coords = np.array([((x + 1) / 2, (y + 1) / 2) for x, y, _ in a] + [(0.5, 0.5)] * 4, np.float32)
    spec.texture_indices = spec.vertex_indices.copy()
    spec.textures = []
    return spec


def quad_a_rays(n=400, seed=0):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(-0.9, 0.9, (4 * n, 2)).astype(np.float32)
    xy = xy[np.abs(xy[:, 0] - xy[:, 1]) > 0.05][:n]
    o = np.column_stack([xy, np.zeros(len(xy), np.float32)])
    s = rng.choice(np.array([0.5, 1.0, 3.7], np.float32), len(xy))
    d = np.column_stack([np.zeros((len(xy), 2), np.float32), s])
    return o, d, xy


def quad_b_rays(n=200, seed=1):
    rng = np.random.default_rng(seed)
    y = rng.uniform(-0.9, 0.9, n).astype(np.float32)
    z = rng.uniform(0.0, 8.0, n).astype(np.float32)
    z = np.where(np.abs(z - 2.0) < 0.05, z + 0.1, z).astype(np.float32)
    z = np.where(np.abs(z - (5 * y + 4)) < 0.05, z + 0.2, z).astype(np.float32)   # off the diagonal
    o = np.column_stack([np.zeros(n, np.float32), y, z])
    d = np.column_stack([rng.uniform(0.3, 2.0, n), np.zeros(n), np.zeros(n)]).astype(np.float32)
    return o, d


def check_quad_a(rec, xy, ulps=2):
    n = len(xy)
    assert np.all(rec["material"][:n] == 0)
    np.testing.assert_array_equal(rec["prim"], np.where(xy[:, 0] > xy[:, 1], 0, 1))
    np.testing.assert_allclose(rec["distance"], 2.0, rtol=ulps * 2.0 ** -23, atol=0)
    np.testing.assert_array_equal(rec["point"][:, :2], xy)
    np.testing.assert_allclose(rec["point"][:, 2], 2.0, rtol=ulps * 2.0 ** -23, atol=0)
    np.testing.assert_array_equal(rec["normal"], np.tile(np.float32([0, 0, 1]), (n, 1)))
    np.testing.assert_allclose(rec["uv"], (xy.astype(np.float64) + 1) / 2, rtol=0, atol=4e-7)


def check_quad_b(rec, o):
    assert np.all(rec["material"] == 1)
    np.testing.assert_array_equal(rec["prim"], np.where(o[:, 2] > o[:, 1] * 5 + 4, 3, 2))
    np.testing.assert_allclose(rec["distance"], 5.0, rtol=2 * 2.0 ** -23, atol=0)
    np.testing.assert_array_equal(rec["point"][:, 1:], o[:, 1:])
    np.testing.assert_array_equal(rec["normal"], np.tile(np.float32([1, 0, 0]), (len(o), 1)))
    np.testing.assert_allclose(rec["uv"], 0.5, rtol=0, atol=1e-7)


def check_oblique(rec, ab):
    tgt = np.column_stack([ab.astype(np.float64), np.full(len(ab), 2.0)])
    np.testing.assert_allclose(rec["point"], tgt, rtol=4 * 2.0 ** -23, atol=4 * 2.0 ** -24)
    np.testing.assert_allclose(rec["distance"], np.linalg.norm(tgt, axis=1), rtol=4 * 2.0 ** -23)
    assert np.all(rec["material"] == 0)


def check_miss(rec):
    assert np.all(np.isinf(rec["distance"]) & (rec["distance"] > 0))
    assert np.all(rec["prim"] == -1) and np.all(rec["material"] == -1)
    assert not rec["uv"].any() and not rec["normal"].any() and not rec["point"].any()
