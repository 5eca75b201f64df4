"""Pins the oracle's instanced query (pyoracle.intersect_instanced, a restatement of
scene.cpp:79-98 and model.cpp:99-126) against a closed form: tracing an instance with
transform T must find the same primitive at the same world distance as tracing the
geometry transformed by T directly (flattened).  A dropped renormalisation, a transposed
matrix or a distance left in object units fails these checks."""
import numpy as np

import pyoracle as O
from crender_amd import scenes


def _xf(angle_deg, axis, scale, trans):
    a = np.deg2rad(angle_deg)
    c, s = np.cos(a), np.sin(a)
    ax = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + s * K + (1 - c) * K @ K
    M = np.eye(4)
    M[:3, :3] = R @ np.diag(scale)
    M[:3, 3] = trans
    return M.T.reshape(-1).astype(np.float32)      # column-major


def _flatten(tri, T):
    M = np.asarray(T, np.float64).reshape(4, 4).T
    p = tri.reshape(-1, 3).astype(np.float64) @ M[:3, :3].T + M[:3, 3]
    return p.astype(np.float32).reshape(-1, 9)


def test_identity_instances_equal_plain_query():
    a = scenes.soup(3000, seed=11).corners().reshape(-1, 9)
    b = scenes.soup(2000, seed=12).corners().reshape(-1, 9) + np.float32(0.3)
    org, dirs = scenes.diffuse_rays(3000, seed=5, lo=-0.2, hi=1.4)
    I = np.eye(4, dtype=np.float32).reshape(-1)
    m, i, p, d, u, v = O.intersect_instanced([a, b], [[I], [I]], org, dirs)
    prim, t, _, _ = O.intersect(np.concatenate([a, b]), org, dirs)
    hit = prim >= 0
    assert np.array_equal(m >= 0, hit)
    gp = np.where(m == 1, p + a.shape[0], p)
    assert np.mean(gp[hit] == prim[hit]) > 0.999
    dn = np.linalg.norm(dirs.astype(np.float64), axis=1)
    np.testing.assert_allclose(d[hit], t[hit] * dn[hit], rtol=2e-5)


def test_transformed_instance_equals_flattened_geometry():
    tri = scenes.soup(4000, seed=13).corners().reshape(-1, 9)
    Ts = [_xf(30, (0, 1, 0), (1, 1, 1), (0.2, 0, 0)), _xf(-70, (1, 1, 0), (2.0, 0.5, 1.5), (0.1, 0.4, -0.3)),
          _xf(10, (0, 0, 1), (-1.0, 1, 1), (1.0, 0.0, 0.5))]          # a mirror (negative determinant)
    org, dirs = scenes.diffuse_rays(4000, seed=6, lo=-0.8, hi=1.8)
    m, i, p, d, u, v = O.intersect_instanced([tri], [Ts], org, dirs)
    flat = np.concatenate([_flatten(tri, T) for T in Ts])
    prim, t, fu, fv = O.intersect(flat, org, dirs)
    hit = prim >= 0
    assert np.array_equal(m >= 0, hit)
    # same instance and triangle (except near-ties), same world distance
    same = (i[hit] * tri.shape[0] + p[hit]) == prim[hit]
    assert same.mean() > 0.998, same.mean()
    dn = np.linalg.norm(dirs.astype(np.float64), axis=1)
    close = np.isclose(d[hit], t[hit] * dn[hit], rtol=1e-4)
    assert close.mean() > 0.999
    np.testing.assert_allclose(u[hit][same], fu[hit][same], atol=2e-3)
