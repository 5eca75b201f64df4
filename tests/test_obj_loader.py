"""OBJ/MTL ingest (crt_obj_load): host code, runs without a GPU.

Pins tinyobjloader's triangulation rules as asset_loader.cpp:182-303 uses them (quads split
on the shorter diagonal, ear clipping for larger polygons, relative indices), the MTL
fields the reference keeps (Kd, map_Kd flipped vertically, smooth, emission 0), the
defaults appended where the reference would index out of bounds, and a round trip of the
seeded Cornell box through OBJ.
"""
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

from crender_amd import CrtError, ObjModel, scenes


def _obj(tmp_path, text, name="m.obj", mtl=None):
    p = tmp_path / name
    p.write_text(text)
    if mtl is not None:
        (tmp_path / "m.mtl").write_text(mtl)
    return p


def test_quad_split_on_shorter_diagonal(tmp_path):
    # |v2 - v0|^2 = 2 < |v3 - v1|^2 = 4.25: [0,1,2],[0,2,3]
    p = _obj(tmp_path, "v 0 0 0\nv 1 0 0\nv 1 1 0\nv -1 1.5 0\nf 1 2 3 4\n")
    m = ObjModel(p)
    assert m.vertex_indices.tolist() == [[0, 1, 2], [0, 2, 3]]
    # mirrored so the other diagonal is shorter: [0,1,3],[1,2,3]
    p = _obj(tmp_path, "v 0 0 0\nv 1 0 0\nv 3 2 0\nv 0 1 0\nf 1 2 3 4\n", name="q2.obj")
    m = ObjModel(p)
    assert m.vertex_indices.tolist() == [[0, 1, 3], [1, 2, 3]]


def test_convex_pentagon_ear_order(tmp_path):
    ang = np.deg2rad(np.arange(5) * 72.0)
    vs = "".join(f"v {np.cos(a):.6f} {np.sin(a):.6f} 0\n" for a in ang)
    m = ObjModel(_obj(tmp_path, vs + "f 1 2 3 4 5\n"))
    assert m.vertex_indices.tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4]]


def _area2(p):
    x, y = p[:, 0], p[:, 1]
    return 0.5 * np.sum(x * np.roll(y, -1) - y * np.roll(x, -1))


def test_concave_polygon_tiles_its_area(tmp_path):
    # L-shaped hexagon in the xz plane (y = 2): triangles must tile it exactly
    poly = np.array([[0, 0], [3, 0], [3, 1], [1, 1], [1, 3], [0, 3]], float)
    vs = "".join(f"v {x} 2 {z}\n" for x, z in poly)
    m = ObjModel(_obj(tmp_path, vs + "f 1 2 3 4 5 6\n"))
    tris = m.vertex_indices
    assert tris.shape == (4, 3)
    tot = 0.0
    for t in tris:
        a = _area2(poly[t])
        assert a > 0          # same orientation as the polygon: no reflex ear was cut
        tot += a
    assert abs(tot - _area2(poly)) < 1e-9


def test_relative_indices_and_missing_material(tmp_path):
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nf -3 -2 -1\nv 0 0 1\nf 1 2 -1\n"
    m = ObjModel(_obj(tmp_path, text))
    assert m.vertex_indices.tolist() == [[0, 1, 2], [0, 1, 3]]
    # no usemtl: one appended default material (white smooth)
    assert len(m.materials) == 1 and m.materials[0]["type"] == 1
    assert m.material_indices.tolist() == [0, 0]
    assert m.texture_coords is None


def test_mtl_colours_and_groups(tmp_path):
    mtl = "newmtl red\nKd 0.8 0.1 0.1\nKe 5 5 5\n\nnewmtl white\nKd 0.7 0.7 0.7\n"
    text = ("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\n"
            "usemtl white\nf 1 2 3\nusemtl red\nf 2 4 3\nusemtl nosuch\nf 1 2 4\n")
    m = ObjModel(_obj(tmp_path, text, mtl=mtl))
    assert len(m.materials) == 3            # red, white, appended default
    assert m.materials[0]["colour"][:3] == pytest.approx((0.8, 0.1, 0.1))
    assert m.materials[0]["emission"] == 0.0   # Ke ignored (asset_loader.cpp:233)
    assert m.materials[1]["colour"][:3] == pytest.approx((0.7, 0.7, 0.7))
    assert m.material_indices.tolist() == [1, 0, 2]


def test_partial_texture_coordinates_get_a_default(tmp_path):
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.25 0.5\nvt 1 0\nf 1/1 2/2 3/1\nf 1 2 3\n"
    m = ObjModel(_obj(tmp_path, text))
    assert m.texture_coords.shape == (3, 2) and m.texture_coords[2].tolist() == [0.0, 0.0]
    assert m.texture_indices.tolist() == [[0, 1, 0], [2, 2, 2]]


def _png(w, h, rgba):
    raw = b"".join(b"\x00" + bytes(rgba[y].reshape(-1)) for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def test_textures_png_and_ppm_flipped(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (3, 4, 4), dtype=np.uint8)          # 4 wide, 3 high
    (tmp_path / "t.png").write_bytes(_png(4, 3, img))
    ppm = rng.integers(0, 256, (2, 5, 3), dtype=np.uint8)
    (tmp_path / "t.ppm").write_bytes(b"P6\n# c\n5 2\n255\n" + ppm.tobytes())
    mtl = "newmtl a\nKd 1 1 1\nmap_Kd t.png\nnewmtl b\nmap_Kd -bm 1 t.ppm\nnewmtl c\nmap_Kd t.png\nnewmtl d\nmap_Kd none.png\n"
    text = "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl a\nf 1 2 3\nusemtl b\nf 1 2 3\nusemtl c\nf 1 2 3\nusemtl d\nf 1 2 3\n"
    m = ObjModel(_obj(tmp_path, text, mtl=mtl))
    assert [x["tex"] for x in m.materials] == [0, 1, 0, -1]   # shared texture loaded once; missing -> none
    assert "none.png" in m.warnings
    np.testing.assert_array_equal(m.textures[0], img[::-1].astype(np.float32) / 255.0)
    exp = np.concatenate([ppm, np.full((2, 5, 1), 255, np.uint8)], axis=2)[::-1]
    np.testing.assert_array_equal(m.textures[1], exp.astype(np.float32) / 255.0)


def test_bad_index_is_an_error(tmp_path):
    with pytest.raises(CrtError):
        ObjModel(_obj(tmp_path, "v 0 0 0\nv 1 0 0\nf 1 2 7\n"))


def test_cornell_round_trip(tmp_path):
    spec = scenes.cornell(1)
    p = tmp_path / "cornell.obj"
    scenes.write_obj(spec, p)
    m = ObjModel(p)
    assert m.name == "cornell"
    np.testing.assert_array_equal(m.vertices[m.vertex_indices].reshape(-1, 9), spec.corners().reshape(-1, 9))
    assert m.material_indices.tolist() == list(np.asarray(spec.material_indices))
    for got, want in zip(m.materials, spec.materials):
        assert got["colour"][:3] == pytest.approx(tuple(want.colour[:3]))


# ---- against the reference's own tinyobjloader (oracle/_ref, built where the reference is) --
REF_TINYOBJ = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "tinyobj_dump"


def _tinyobj(path):
    """Triangles (vertex indices, texcoord indices, material) of tinyobjloader's reading."""
    out = subprocess.run([str(REF_TINYOBJ), str(path)], capture_output=True, text=True, check=True).stdout
    tris, tcs, mats = [], [], []
    for line in out.splitlines():
        w = line.split()
        assert w[0] != "n", f"tinyobj left a face of {w[1]} vertices untriangulated"
        if w[0] == "f":
            tris.append([int(x) for x in w[1:4]])
            tcs.append([int(x) for x in w[4:7]])
            mats.append(int(w[7]))
    return np.array(tris, np.int64).reshape(-1, 3), np.array(tcs, np.int64).reshape(-1, 3), np.array(mats)


def _random_polygon(rng, n):
    """A simple (star-shaped) polygon with n vertices: sorted angles, random radii."""
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    rad = rng.uniform(0.3, 1.0, n)
    return np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)


needs_ref = pytest.mark.skipif(not REF_TINYOBJ.exists(), reason="oracle/_ref/tinyobj_dump not built "
                               "(the reference tree is absent)")


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_triangulation_matches_reference_tinyobj(tmp_path, seed):
    """Random star-shaped concave polygons of 4-12 vertices in random planes, quads, relative
    indices, texcoords and two materials: the loader's triangles and primID order equal
    tinyobjloader's (tinobj.h:1393-1714) face for face."""
    rng = np.random.default_rng(seed)
    lines, nv, nt = ["mtllib m.mtl"], 0, 0
    for f in range(40):
        n = int(rng.integers(3, 13))
        poly = _random_polygon(rng, n)
        # embed in a random plane (rotation), so the dominant-axis projection varies
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        pts = np.concatenate([poly, rng.normal(scale=1e-3, size=(n, 1)) * 0], 1) @ q.T + rng.normal(size=3)
        for p in pts:
            lines.append(f"v {p[0]:.7f} {p[1]:.7f} {p[2]:.7f}")
        for p in poly:
            lines.append(f"vt {0.5 + 0.5 * p[0]:.6f} {0.5 + 0.5 * p[1]:.6f}")
        if f % 7 == 0:
            lines.append(f"usemtl m{f % 2}")
        if f % 3 == 0:   # relative indices
            idx = " ".join(f"{k - n}/{k - n}" for k in range(n))
        else:
            idx = " ".join(f"{nv + k + 1}/{nt + k + 1}" for k in range(n))
        lines.append("f " + idx)
        nv += n
        nt += n
    mtl = "newmtl m0\nKd 0.8 0.2 0.2\n\nnewmtl m1\nKd 0.2 0.8 0.2\n"
    p = _obj(tmp_path, "\n".join(lines) + "\n", mtl=mtl)
    ref_v, ref_t, ref_m = _tinyobj(p)
    m = ObjModel(p)
    assert m.vertex_indices.shape == ref_v.shape
    np.testing.assert_array_equal(m.vertex_indices, ref_v)
    np.testing.assert_array_equal(m.texture_indices, ref_t)
    # faces before the first usemtl have material -1 in tinyobj; the loader appends a default
    got_m = np.asarray(m.material_indices, np.int64)
    named = ref_m >= 0
    np.testing.assert_array_equal(got_m[named], ref_m[named])


@needs_ref
def test_cornell_obj_matches_reference_tinyobj(tmp_path):
    spec = scenes.cornell(3)
    path = tmp_path / "cornell.obj"
    scenes.write_obj(spec, path)
    ref_v, _, ref_m = _tinyobj(path)
    m = ObjModel(path)
    np.testing.assert_array_equal(m.vertex_indices, ref_v)
    np.testing.assert_array_equal(np.asarray(m.material_indices, np.int64), ref_m)
