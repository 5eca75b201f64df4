"""Pins the oracle's intersection records (scene::cast_ray, scene.cpp:79-98; model.cpp:27-47)
with closed forms on two axis-aligned quads, none of which restates the oracle's code:

  * quad A in the plane z = 2 over [-1, 1]^2 with uv = ((x + 1) / 2, (y + 1) / 2) at its
    corners, material 0: a ray (x, y, 0) + s (0, 0, 1) for any scale s hits it at distance 2,
    point (x, y, 2), normal (0, 0, 1) (the winding's (v1 - v0) x (v2 - v0)), uv the affine
    map, and the triangle on its side of the diagonal x = y;
  * quad B in the plane x = 5, material 1, constant uv 0.5: rays along +x hit it at 5;
  * oblique rays from the origin to (a, b, 2): the point (a, b, 2), distance |(a, b, 2)|;
  * rays pointing away miss: distance +inf, IDs -1, zeros elsewhere;
  * tnear = 1e-5 (model.cpp:22): an origin on the plane or 5e-6 below it sees nothing, one
    1e-3 below it hits at 1e-3.
"""
import numpy as np
import pytest

from _records import check_miss, check_oblique, check_quad_a, check_quad_b, quad_a_rays, quad_b_rays, quads_spec


def _oracle_rec(oracle, spec, o, d):
    r = oracle.cast_records(spec, o, d, trav=oracle.BRUTE32)
    return {"distance": r["distance"], "prim": r["prim_id"], "material": r["material"], "uv": r["uv"],
            "normal": r["normal"], "point": r["point"]}


def test_quad_a_closed_forms(oracle):
    spec = quads_spec()
    o, d, xy = quad_a_rays()
    check_quad_a(_oracle_rec(oracle, spec, o, d), xy)


def test_quad_b_closed_forms(oracle):
    spec = quads_spec()
    o, d = quad_b_rays()
    check_quad_b(_oracle_rec(oracle, spec, o, d), o)


def test_oblique_rays(oracle):
    ab = np.random.default_rng(2).uniform(-0.9, 0.9, (300, 2)).astype(np.float32)
    ab = ab[np.abs(ab[:, 0] - ab[:, 1]) > 0.01]
    d = np.column_stack([ab, np.full(len(ab), 2.0, np.float32)])
    check_oblique(_oracle_rec(oracle, quads_spec(), np.zeros_like(d), d), ab)


def test_misses(oracle):
    o = np.zeros((3, 3), np.float32)
    d = np.array([(0, 0, -1), (-1, 0, 0), (0, 1, 0)], np.float32)
    check_miss(_oracle_rec(oracle, quads_spec(), o, d))


@pytest.mark.parametrize("below,hits", [(0.0, False), (5e-6, False), (1e-3, True)])
def test_tnear(oracle, below, hits):
    o = np.array([(0.3, -0.2, 2.0 - below)], np.float32)
    d = np.array([(0, 0, 1)], np.float32)
    r = _oracle_rec(oracle, quads_spec(), o, d)
    if hits:
        assert r["prim"][0] == 0
        np.testing.assert_allclose(r["distance"][0], np.float32(2.0) - np.float32(2.0 - below), rtol=1e-6)
    else:
        check_miss(r)
