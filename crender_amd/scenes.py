"""Seeded synthetic scenes for the path tracer's workloads.

Each generator returns a :class:`SceneSpec`: indexed geometry in the shape of CRender's
``asset_loader::model_data`` (asset_loader.h:16-30) plus materials, sky, sun and camera.
All randomness comes from :mod:`crender_amd.philox`, so a scene is a pure function of
its seed.

Workloads:
  * ``cornell``  -- 36-triangle Cornell box: 5 walls, two boxes rotated by seeded angles
    in [0, 30] degrees, one emissive quad.  Rendered at 256x256, 16 spp, depth 8,
    Russian roulette after bounce 3.
  * ``soup``     -- random triangle soup: centroids uniform in the unit cube, vertices
    offset from the centroid by N(0, 0.005^2 I).  1M triangles at 1920x1080.
  * ``terrain``  -- fBm height field plus icospheres with diffuse / GGX / dielectric
    materials under a sun and constant sky.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .philox import Stream

SMOOTH, METAL, GLASS = 1, 0, 2


@dataclass
class MaterialSpec:
    type: int = SMOOTH
    ior: float = 1.5
    roughness: float = 0.5
    reflectiveness: float = 1.0
    emission: float = 0.0
    colour: tuple = (1.0, 1.0, 1.0, 1.0)
    tex: int = -1


@dataclass
class SceneSpec:
    name: str
    vertices: np.ndarray                 # (V, 3) float32
    vertex_indices: np.ndarray           # (3T,) uint32
    material_indices: np.ndarray         # (T,) uint32
    materials: list
    texture_coords: np.ndarray | None = None   # (U, 2) float32
    texture_indices: np.ndarray | None = None  # (3T,) uint32
    textures: list = field(default_factory=list)  # list of (h, w, 4) float32
    sky: np.ndarray | None = None        # (h, w, 4) float32 or None (black)
    sky_rotation: tuple = (0.0, 0.0)     # added to the equirect (u, v) (scene.cpp:62-77)
    sun_enabled: bool = False
    sun: dict | None = None
    camera: dict = field(default_factory=dict)
    res: tuple = (256, 256)
    spp: int = 1
    max_bounces: int = 8
    rr_start: int = 0
    mode: int = 0                        # 0 compat, 1 pbr

    @property
    def ntri(self) -> int:
        return int(self.material_indices.shape[0])

    def corners(self) -> np.ndarray:
        """Expanded (T, 9) corners in primID order (registry.cpp:13-24)."""
        return self.vertices[self.vertex_indices].reshape(-1, 9).astype(np.float32)

    def corner_uv(self) -> np.ndarray | None:
        if self.texture_coords is None:
            return None
        return self.texture_coords[self.texture_indices].reshape(-1, 6).astype(np.float32)


def _quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); winding gives normal (p1-p0)x(p2-p0)."""
    return [(p0, p1, p2), (p0, p2, p3)]


def _box(size, angle_deg, pos):
    """12 outward-facing triangles of an axis-aligned box rotated about +y."""
    sx, sy, sz = size
    a = math.radians(angle_deg)
    c, s = math.cos(a), math.sin(a)

    def xf(x, y, z):
        return (pos[0] + c * x + s * z, pos[1] + y, pos[2] - s * x + c * z)

    hx, hz = sx / 2, sz / 2
    v = {
        "lbf": xf(-hx, 0, -hz), "rbf": xf(hx, 0, -hz), "rbb": xf(hx, 0, hz), "lbb": xf(-hx, 0, hz),
        "ltf": xf(-hx, sy, -hz), "rtf": xf(hx, sy, -hz), "rtb": xf(hx, sy, hz), "ltb": xf(-hx, sy, hz),
    }
    tris = []
    tris += _quad(v["lbf"], v["rbf"], v["rbb"], v["lbb"])  # bottom (normal -y)
    tris += _quad(v["ltf"], v["ltb"], v["rtb"], v["rtf"])  # top (+y)
    tris += _quad(v["lbf"], v["ltf"], v["rtf"], v["rbf"])  # front (-z)
    tris += _quad(v["lbb"], v["rbb"], v["rtb"], v["ltb"])  # back (+z)
    tris += _quad(v["lbf"], v["lbb"], v["ltb"], v["ltf"])  # left (-x)
    tris += _quad(v["rbf"], v["rtf"], v["rtb"], v["rbb"])  # right (+x)
    return tris


def _from_tris(tris, mats_per_tri):
    verts = np.asarray([p for t in tris for p in t], dtype=np.float32)
    idx = np.arange(len(verts), dtype=np.uint32)
    return verts, idx, np.asarray(mats_per_tri, dtype=np.uint32)


def camera_dict(position, rotation=(0.0, 0.0, 0.0), fov=75.0, mode=0, scale=1.0, matrix=None):
    """Camera with the _cached_matrix of camera.cpp:54-68 (identity rotation here unless given)."""
    if matrix is None:
        matrix = camera_matrix(position, rotation)
    return {"position": tuple(map(float, position)), "rotation": tuple(map(float, rotation)),
            "fov": float(fov), "mode": int(mode), "scale": float(scale),
            "matrix": np.asarray(matrix, dtype=np.float32).reshape(16)}


def _rot(axis, deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    x, y, z = axis
    t = 1 - c
    return np.array([[c + t * x * x, t * x * y - s * z, t * x * z + s * y, 0],
                     [t * x * y + s * z, c + t * y * y, t * y * z - s * x, 0],
                     [t * x * z - s * y, t * y * z + s * x, c + t * z * z, 0],
                     [0, 0, 0, 1]], dtype=np.float64)


def camera_matrix(position, rotation):
    """T(position) * Ry(rot.x) * Rx(rot.y) * Rz(rot.z), column-major flat (camera.cpp:54-68)."""
    m = np.eye(4)
    m[:3, 3] = position
    m = m @ _rot((0, 1, 0), rotation[0]) @ _rot((1, 0, 0), rotation[1]) @ _rot((0, 0, 1), rotation[2])
    return m.T.reshape(16).astype(np.float32)   # column-major


def constant_sky(value=1.0):
    return np.full((1, 1, 4), value, dtype=np.float32)


def cornell(seed: int = 1) -> SceneSpec:
    """Cornell box; normals of walls and light face the interior, boxes face outward."""
    st = Stream(seed, stream=1)
    ang = st.uniform(2, dims=1)[:, 0] * 30.0
    red, green, white, light = 0, 1, 2, 3
    tris, mats = [], []

    def add(ts, m):
        tris.extend(ts)
        mats.extend([m] * len(ts))

    add(_quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)), white)     # floor, +y
    add(_quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)), white)     # ceiling, -y
    add(_quad((-1, 0, 1), (-1, 2, 1), (1, 2, 1), (1, 0, 1)), white)       # back, -z
    add(_quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1)), red)     # left, +x
    add(_quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)), green)       # right, -x
    add(_box((0.6, 1.2, 0.6), float(ang[0]), (-0.35, 0.0, 0.3)), white)
    add(_box((0.6, 0.6, 0.6), float(ang[1]), (0.35, 0.0, -0.3)), white)
    add(_quad((-0.25, 1.98, -0.25), (0.25, 1.98, -0.25), (0.25, 1.98, 0.25), (-0.25, 1.98, 0.25)), light)  # -y
    verts, idx, mi = _from_tris(tris, mats)
    materials = [
        MaterialSpec(colour=(0.63, 0.065, 0.05, 1.0)),
        MaterialSpec(colour=(0.14, 0.45, 0.091, 1.0)),
        MaterialSpec(colour=(0.725, 0.71, 0.68, 1.0)),
        MaterialSpec(colour=(1.0, 1.0, 1.0, 1.0), emission=15.0),
    ]
    uv = np.zeros((len(verts), 2), dtype=np.float32)
    return SceneSpec(name=f"cornell-{seed}", vertices=verts, vertex_indices=idx, material_indices=mi,
                     materials=materials, texture_coords=uv, texture_indices=idx.copy(),
                     sky=None, sun_enabled=False, camera=camera_dict((0.0, 1.0, -3.4), fov=40.0),
                     res=(256, 256), spp=16, max_bounces=8, rr_start=3, mode=0)


def soup(n: int = 1_000_000, seed: int = 2, res=(1920, 1080), sigma: float = 0.005) -> SceneSpec:
    """Random triangle soup; sky constant 1 so the image is not black (the sun is off)."""
    st = Stream(seed, stream=2)
    cen = st.uniform(n, dims=3).astype(np.float64)
    off = Stream(seed, stream=3).normal(3 * n, dims=3) * sigma
    verts = (np.repeat(cen, 3, axis=0) + off).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32)
    mi = np.zeros(n, dtype=np.uint32)
    return SceneSpec(name=f"soup-{n}-{seed}", vertices=verts, vertex_indices=idx, material_indices=mi,
                     materials=[MaterialSpec(colour=(0.8, 0.8, 0.8, 1.0))], sky=constant_sky(1.0),
                     sun_enabled=False, camera=camera_dict((0.5, 0.5, -1.0), fov=60.0),
                     res=tuple(res), spp=1, max_bounces=4, rr_start=0, mode=0)


def furnace(albedo: float = 0.5, emission: float = 1.0) -> SceneSpec:
    """Closed cube, every face emissive with the same albedo: L = E*a / (1 - a) inside
    (compat: emission is added after the albedo multiply, renderer.cpp:310-311)."""
    tris = []
    p = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
    faces = [(0, 1, 2, 3), (4, 7, 6, 5), (0, 3, 7, 4), (1, 5, 6, 2), (0, 4, 5, 1), (3, 2, 6, 7)]
    for f in faces:   # inward normals
        tris += _quad(p[f[0]], p[f[1]], p[f[2]], p[f[3]])
    verts, idx, mi = _from_tris(tris, [0] * len(tris))
    return SceneSpec(name="furnace", vertices=verts, vertex_indices=idx, material_indices=mi,
                     materials=[MaterialSpec(colour=(albedo, albedo, albedo, 1.0), emission=emission)],
                     sky=None, sun_enabled=False, camera=camera_dict((0.0, 0.0, 0.0), fov=60.0),
                     res=(32, 32), spp=1, max_bounces=64, rr_start=0, mode=0)


def rays_from_camera(spec: SceneSpec, res, seed: int = 7, jitter: bool = True):
    """Primary rays (SoA float32) of a pinhole camera, for traversal parity tests."""
    W, H = res
    xs, ys = np.meshgrid(np.arange(W, dtype=np.float32), np.arange(H, dtype=np.float32))
    xs = xs.reshape(-1)
    ys = ys.reshape(-1)
    if jitter:
        u = Stream(seed, stream=9).uniform(W * H, dims=2)
        xs = xs + u[:, 0]
        ys = ys + u[:, 1]
    else:
        xs = xs + 0.5
        ys = ys + 0.5
    cam = spec.camera
    m = np.asarray(cam["matrix"], dtype=np.float64).reshape(4, 4).T
    aspect = W / H
    u = (2.0 * xs / W - 1.0) * aspect
    v = 2.0 * ys / H - 1.0
    w = 1.0 / math.tan(0.5 * math.radians(cam["fov"]))
    d = (m[:3, :3] @ np.stack([u, v, np.full_like(u, w)]).astype(np.float64)).T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(np.asarray(cam["position"], dtype=np.float32), d.shape)
    return o.astype(np.float32), d.astype(np.float32)


def diffuse_rays(n: int, seed: int, lo=0.0, hi=1.0):
    """Rays with origins uniform in a box and uniformly random directions."""
    u = Stream(seed, stream=10).uniform(n, dims=6).astype(np.float64)
    o = lo + (hi - lo) * u[:, :3]
    z = 2.0 * u[:, 3] - 1.0
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    ph = 2.0 * np.pi * u[:, 4]
    d = np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=1)
    return o.astype(np.float32), d.astype(np.float32)


def sunlit(seed: int = 4, mode: int = 0) -> SceneSpec:
    """Ground quad and four boxes (diffuse, GGX metal, mirror, glass) under the default sun
    and a constant sky: exercises every material branch and the sun NEE."""
    st = Stream(seed, stream=4)
    ang = st.uniform(4, dims=1)[:, 0] * 30.0
    tris, mats = [], []

    def add(ts, m):
        tris.extend(ts)
        mats.extend([m] * len(ts))

    add(_quad((-4, 0, -4), (-4, 0, 4), (4, 0, 4), (4, 0, -4)), 0)
    add(_box((0.8, 0.8, 0.8), float(ang[0]), (-1.2, 0.0, 0.5)), 1)
    add(_box((0.8, 1.2, 0.8), float(ang[1]), (0.0, 0.0, 1.0)), 2)
    add(_box((0.6, 0.6, 0.6), float(ang[2]), (1.2, 0.0, 0.2)), 3)
    add(_box((0.5, 0.9, 0.5), float(ang[3]), (0.4, 0.0, -0.6)), 4)
    verts, idx, mi = _from_tris(tris, mats)
    materials = [
        MaterialSpec(colour=(0.6, 0.6, 0.6, 1.0)),
        MaterialSpec(colour=(0.8, 0.3, 0.2, 1.0)),
        MaterialSpec(type=METAL, roughness=0.3, colour=(0.9, 0.8, 0.5, 1.0)),
        MaterialSpec(type=METAL, roughness=0.0, colour=(0.95, 0.95, 0.95, 1.0), reflectiveness=0.9),
        MaterialSpec(type=GLASS, ior=1.5, colour=(0.95, 0.95, 1.0, 1.0)),
    ]
    cam = camera_dict((0.0, 1.5, -4.0), rotation=(0.0, -12.0, 0.0), fov=50.0)
    return SceneSpec(name=f"sunlit-{seed}", vertices=verts, vertex_indices=idx, material_indices=mi,
                     materials=materials, sky=constant_sky(0.4), sun_enabled=True, sun=None, camera=cam,
                     res=(96, 64), spp=4, max_bounces=6, rr_start=3, mode=mode)


def cutout(seed: int = 5, mode: int = 0) -> SceneSpec:
    """Textures, alpha cut-outs and a rotated equirect sky under the sun: a ground quad with a
    5x3 checker texture tiled over uv in [-1.3, 2.7] (negative coordinates wrap, image.h:
    104-109), a fence quad whose 4x4 texture has transparent columns (colour.w == 0, so camera
    rays and sun shadow rays pass through, renderer.cpp:294-301,330-345), a fully transparent
    quad in front of part of the view, diffuse / glass / metal boxes, and a 16x8 sky image
    with a rotation that takes the lookup below 0 (scene.cpp:62-77)."""
    st = Stream(seed, stream=7)
    ang = st.uniform(3, dims=1)[:, 0] * 30.0
    tris, mats, uvs = [], [], []

    def add(ts, m, uv=None):
        tris.extend(ts)
        mats.extend([m] * len(ts))
        uvs.extend(uv if uv is not None else [((0, 0), (0, 0), (0, 0))] * len(ts))

    g = [(-4, 0, -4), (-4, 0, 4), (4, 0, 4), (4, 0, -4)]
    gu = [(-1.3, -0.7), (-1.3, 2.7), (2.7, 2.7), (2.7, -0.7)]
    add(_quad(*g), 0, [(gu[0], gu[1], gu[2]), (gu[0], gu[2], gu[3])])
    f = [(-1.5, 0, 1.6), (-1.5, 1.4, 1.6), (1.5, 1.4, 1.6), (1.5, 0, 1.6)]       # fence, facing -z
    fu = [(0, 0), (0, 1), (2, 1), (2, 0)]
    add(_quad(*f), 1, [(fu[0], fu[1], fu[2]), (fu[0], fu[2], fu[3])])
    add(_quad((-0.9, 0.2, -1.5), (-0.9, 1.0, -1.5), (-0.1, 1.0, -1.5), (-0.1, 0.2, -1.5)), 2)   # clear pane
    add(_box((0.6, 0.6, 0.6), float(ang[0]), (-1.0, 0.0, 0.4)), 3)
    add(_box((0.6, 0.9, 0.6), float(ang[1]), (0.2, 0.0, 0.7)), 4)
    add(_box((0.5, 0.5, 0.5), float(ang[2]), (1.1, 0.0, 0.1)), 5)
    verts, idx, mi = _from_tris(tris, mats)
    uv = np.asarray([c for t in uvs for c in t], dtype=np.float32)
    checker = np.zeros((3, 5, 4), np.float32)
    for y in range(3):
        for x in range(5):
            v = 0.25 + 0.5 * ((x + y) % 2)
            checker[y, x] = (v, 0.6 * v + 0.1 * x, 0.3 + 0.1 * y, 1.0)
    fence = np.zeros((4, 4, 4), np.float32)
    for y in range(4):
        for x in range(4):
            fence[y, x] = (0.7, 0.5, 0.3, 0.0 if x % 2 else 1.0)
    sky = np.zeros((8, 16, 4), np.float32)
    for y in range(8):
        for x in range(16):
            sky[y, x] = (0.2 + 0.05 * x, 0.3 + 0.08 * y, 0.6 - 0.05 * y + 0.2 * ((x + y) % 2), 1.0)
    materials = [
        MaterialSpec(colour=(1.0, 1.0, 1.0, 1.0), tex=0),
        MaterialSpec(colour=(1.0, 1.0, 1.0, 1.0), tex=1),
        MaterialSpec(colour=(0.2, 0.9, 0.2, 0.0)),
        MaterialSpec(colour=(0.8, 0.3, 0.2, 1.0)),
        MaterialSpec(type=GLASS, ior=1.5, colour=(0.95, 0.95, 1.0, 1.0)),
        MaterialSpec(type=METAL, roughness=0.3, reflectiveness=0.9, colour=(0.9, 0.8, 0.5, 1.0)),
    ]
    cam = camera_dict((0.0, 1.4, -3.5), rotation=(0.0, -14.0, 0.0), fov=55.0)
    return SceneSpec(name=f"cutout-{seed}", vertices=verts, vertex_indices=idx, material_indices=mi,
                     materials=materials, texture_coords=uv, texture_indices=idx.copy(),
                     textures=[checker, fence], sky=sky, sky_rotation=(0.13, -0.21), sun_enabled=True,
                     sun={"direction": (0.3, -0.6, -0.75), "colour": (1.0, 0.95, 0.85), "intensity": 60.0},
                     camera=cam, res=(96, 64), spp=4, max_bounces=6, rr_start=3, mode=mode)


def _value_noise(x, z, freq, st: Stream, octave: int):
    """Value noise on a (freq+1)^2 lattice of Philox uniforms in [-1, 1], smoothstep-blended."""
    n = freq + 1
    lat = st.uniform(n * n, block=octave, dims=1)[:, 0].astype(np.float64).reshape(n, n) * 2.0 - 1.0
    gx, gz = x * freq, z * freq
    i0 = np.clip(np.floor(gx).astype(np.int64), 0, freq - 1)
    j0 = np.clip(np.floor(gz).astype(np.int64), 0, freq - 1)
    fx, fz = gx - i0, gz - j0
    sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
    a = lat[i0, j0] + (lat[i0 + 1, j0] - lat[i0, j0]) * sx
    b = lat[i0, j0 + 1] + (lat[i0 + 1, j0 + 1] - lat[i0, j0 + 1]) * sx
    return a + (b - a) * sz


def fbm(x, z, st: Stream, octaves=8, lacunarity=2.0, gain=0.5, base=4):
    total, amp, norm = 0.0, 1.0, 0.0
    for o in range(octaves):
        total = total + amp * _value_noise(x, z, int(base * lacunarity ** o), st, o)
        norm += amp
        amp *= gain
    return total / norm


def icosphere(subdiv: int = 2):
    """Unit icosphere: 20 * 4^subdiv outward-facing triangles."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(verts, np.float64), np.array(f, np.int64)


def terrain(grid: int = 2048, spheres: int = 20_000, seed: int = 3, res=(1920, 1080)) -> SceneSpec:
    """fBm terrain (8 octaves, amplitude 0.2, lacunarity 2, gain 0.5) on a grid x grid
    vertex lattice over [0,1]^2, plus icospheres (subdivision 2, 320 triangles) with radius
    log-uniform in [0.002, 0.05] resting on it.  Materials: terrain diffuse; spheres 70%
    diffuse, 20% GGX metal (roughness U[0.05, 0.5]), 10% dielectric (IOR 1.5).  Sun and a
    constant sky; pbr estimator."""
    st = Stream(seed, stream=5)
    g = np.linspace(0.0, 1.0, grid)
    X, Z = np.meshgrid(g, g, indexing="ij")
    Hh = 0.2 * fbm(X, Z, st)
    tv = np.stack([X, Hh, Z], -1).reshape(-1, 3)
    i = np.arange(grid - 1)
    I, J = np.meshgrid(i, i, indexing="ij")
    a = (I * grid + J).reshape(-1)
    b = a + grid
    c = a + 1
    d = b + 1
    # two triangles per cell with +y normals: (a, c, b) and (c, d, b)
    tfaces = np.stack([np.stack([a, c, b], 1), np.stack([c, d, b], 1)], 1).reshape(-1, 3)
    sv, sf = icosphere(2)
    u = Stream(seed, stream=6).uniform(spheres, dims=8).astype(np.float64)
    rad = np.exp(np.log(0.002) + u[:, 0] * (np.log(0.05) - np.log(0.002)))
    px, pz = 0.05 + 0.9 * u[:, 1], 0.05 + 0.9 * u[:, 2]
    py = 0.2 * fbm(px, pz, st) + rad
    centres = np.stack([px, py, pz], 1)
    sverts = (sv[None, :, :] * rad[:, None, None] + centres[:, None, :]).reshape(-1, 3)
    sfaces = (sf[None, :, :] + (np.arange(spheres) * len(sv))[:, None, None]).reshape(-1, 3) + len(tv)
    verts = np.concatenate([tv, sverts]).astype(np.float32)
    faces = np.concatenate([tfaces, sfaces]).astype(np.uint32)
    mats = [MaterialSpec(colour=(0.45, 0.42, 0.35, 1.0))]
    kind = u[:, 3]
    for k in range(spheres):
        col = (0.2 + 0.7 * u[k, 4], 0.2 + 0.7 * u[k, 5], 0.2 + 0.7 * u[k, 6], 1.0)
        if kind[k] < 0.7:
            mats.append(MaterialSpec(colour=col))
        elif kind[k] < 0.9:
            mats.append(MaterialSpec(type=METAL, roughness=0.05 + 0.45 * u[k, 7], colour=col))
        else:
            mats.append(MaterialSpec(type=GLASS, ior=1.5, colour=(0.95, 0.95, 0.95, 1.0)))
    mi = np.concatenate([np.zeros(len(tfaces), np.uint32),
                         np.repeat(np.arange(1, spheres + 1, dtype=np.uint32), len(sf))])
    sky = np.array([[[0.45, 0.55, 0.75, 1.0]]], np.float32)
    cam = camera_dict((0.5, 0.32, -0.25), rotation=(0.0, 22.0, 0.0), fov=55.0)   # pitched down: horizon in frame
    return SceneSpec(name=f"terrain-{grid}-{spheres}-{seed}", vertices=verts, vertex_indices=faces.reshape(-1),
                     material_indices=mi, materials=mats, sky=sky, sun_enabled=True, sun=None, camera=cam,
                     res=tuple(res), spp=64, max_bounces=8, rr_start=3, mode=1)


def write_obj(spec: SceneSpec, path) -> None:
    """Writes a scene's model as Wavefront OBJ + MTL (one `usemtl` group per material, Kd =
    colour).  OBJ carries no emission: the reference's loader sets every material's emission
    to 0 (asset_loader.cpp:233), so emitters are restored with Scene.set_materials."""
    from pathlib import Path

    path = Path(path)
    mtl = path.with_suffix(".mtl")
    with open(mtl, "w") as f:
        for i, m in enumerate(spec.materials):
            c = m.colour
            f.write(f"newmtl m{i}\nKd {c[0]:.9g} {c[1]:.9g} {c[2]:.9g}\n\n")
    tris = np.asarray(spec.vertex_indices, dtype=np.int64).reshape(-1, 3)
    mi = np.asarray(spec.material_indices, dtype=np.int64)
    with open(path, "w") as f:
        f.write(f"# {spec.name}\nmtllib {mtl.name}\n")
        for v in np.asarray(spec.vertices, dtype=np.float32):
            f.write(f"v {v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n")
        cur = None
        for t, m in zip(tris, mi):
            if m != cur:
                f.write(f"usemtl m{m}\n")
                cur = m
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


def batch_suite(n: int = 64, seed: int = 7):
    """Jobs of the batched-rendering workload: `n` independent scenes, alternately triangle
    soup and terrain (with icospheres), triangle counts log-uniform in [1k, 500k],
    resolutions uniform in {128^2, 256^2, 512^2, 1024^2}, spp uniform in {1, 4, 16, 64}.
    Returns a list of (SceneSpec, width, height, spp)."""
    return [(batch_job_scene(p), p["res"], p["res"], p["spp"]) for p in batch_suite_params(n, seed)]


def batch_suite_params(n: int = 64, seed: int = 7):
    """The draws of :func:`batch_suite` without building the scenes: a list of dicts with
    kind ('soup' / 'terrain'), tris, res, spp and the scene seed."""
    u = Stream(seed, stream=20).uniform(n, dims=4).astype(np.float64)
    out = []
    for k in range(n):
        tris = int(np.exp(np.log(1e3) + u[k, 0] * (np.log(5e5) - np.log(1e3))))
        res = (128, 256, 512, 1024)[min(3, int(u[k, 1] * 4))]
        spp = (1, 4, 16, 64)[min(3, int(u[k, 2] * 4))]
        out.append({"index": k, "kind": "soup" if k % 2 == 0 else "terrain", "tris": tris, "res": res, "spp": spp,
                    "seed": seed * 1000 + k})
    return out


def batch_job_scene(p) -> SceneSpec:
    """Scene of one batch job drawn by :func:`batch_suite_params`."""
    res = (p["res"], p["res"])
    if p["kind"] == "soup":
        return soup(p["tris"], seed=p["seed"], res=res)
    spheres = max(1, p["tris"] // 3200)                  # ~10% of the triangles in spheres
    grid = max(2, int(np.sqrt(max(p["tris"] - 320 * spheres, 2) / 2)) + 1)
    return terrain(grid=grid, spheres=spheres, seed=p["seed"], res=res)
