"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker or the timed CPU baseline, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBPATH = HERE / "liboracle.so"

COMPAT, PBR = 0, 1
BRUTE64, BRUTE32, BVH32, BVH64 = 0, 1, 2, 3


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("ior", C.c_float), ("roughness", C.c_float),
                ("reflectiveness", C.c_float), ("emission", C.c_float), ("colour", C.c_float * 4),
                ("tex", C.c_int32)]


class Image(C.Structure):
    _fields_ = [("rgba", C.POINTER(C.c_float)), ("w", C.c_uint32), ("h", C.c_uint32)]


class Sun(C.Structure):
    _fields_ = [("size", C.c_float), ("intensity", C.c_float), ("dir", C.c_float * 3), ("colour", C.c_float * 3)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("fov", C.c_float), ("matrix", C.c_float * 16), ("mode", C.c_int32),
                ("scale", C.c_float)]


class Scene(C.Structure):
    _fields_ = [("tri", C.POINTER(C.c_float)), ("uv", C.POINTER(C.c_float)), ("mat_idx", C.POINTER(C.c_uint32)),
                ("ntri", C.c_uint32), ("mats", C.POINTER(Material)), ("nmat", C.c_uint32),
                ("textures", C.POINTER(Image)), ("ntex", C.c_uint32), ("sky", Image), ("sky_rot", C.c_float * 2),
                ("sun", Sun), ("sun_enabled", C.c_int32), ("cam", Camera), ("bvh", C.c_void_p), ("trav", C.c_int32)]


class Settings(C.Structure):
    _fields_ = [("res_x", C.c_uint32), ("res_y", C.c_uint32), ("max_bounces", C.c_uint32), ("spp", C.c_uint32),
                ("sample_offset", C.c_uint32), ("rr_start", C.c_uint32), ("mode", C.c_uint32), ("seed", C.c_uint32),
                ("aspect", C.c_float), ("row_begin", C.c_uint32), ("row_end", C.c_uint32), ("nthreads", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("ext_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("ref_segments", C.c_uint64),
                ("seconds", C.c_double)]


class ScatterOut(C.Structure):
    _fields_ = [("o", C.c_float * 3), ("d", C.c_float * 3), ("albedo", C.c_float * 3), ("colour", C.c_float * 4),
                ("nee_n", C.c_float * 3), ("emission", C.c_float), ("flags", C.c_int32)]


SCATTER_ALPHA, SCATTER_DELTA, SCATTER_ABSORBED = 1, 2, 4

_L = None
_LIBPATH = LIBPATH


def use_library(path) -> None:
    """Binds this module to another build of the oracle's sources (the CPU baseline's
    optimised build); must be called before the first use."""
    global _LIBPATH
    if _L is not None and Path(path) != _LIBPATH:
        raise RuntimeError("the oracle library is already loaded")
    _LIBPATH = Path(path)


def lib() -> C.CDLL:
    global _L
    if _L is None:
        if not _LIBPATH.exists():
            raise ImportError(f"{_LIBPATH} missing: build with `make -C oracle`")
        L = C.CDLL(str(_LIBPATH))
        fp, up, vp = C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.c_void_p
        L.orc_philox4x32_10.argtypes = [up, up, up]
        L.orc_philox_batch.argtypes = [up, C.c_uint32, up, up]
        L.orc_u01.restype = C.c_float
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_sphere.argtypes = [C.c_float, C.c_float, fp]
        L.orc_hemp_cos.argtypes = [fp, C.c_float, C.c_float, fp]
        L.orc_build_local.argtypes = [fp, fp, fp, fp]
        L.orc_map_to_solid_angle.argtypes = [C.c_float, C.c_float, C.c_float, fp]
        L.orc_solid_angle_pdf.restype = C.c_float
        L.orc_solid_angle_pdf.argtypes = [C.c_float]
        L.orc_tri_watertight32.restype = C.c_int
        L.orc_tri_watertight32.argtypes = [fp, fp, fp, fp, fp, C.c_float, C.c_float, fp, fp, fp]
        L.orc_intersect_batch.argtypes = [fp, C.c_uint32, vp, C.c_int32] + [fp] * 8 + [C.c_uint32, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
        L.orc_bvh_build.restype = vp
        L.orc_bvh_build.argtypes = [fp, C.c_uint32, C.c_uint32]
        L.orc_bvh_free.argtypes = [vp]
        L.orc_bvh_validate.restype = C.c_int
        L.orc_bvh_validate.argtypes = [vp, fp, C.c_uint32]
        L.orc_bvh_node_count.restype = C.c_uint32
        L.orc_bvh_node_count.argtypes = [vp]
        L.orc_bvh_sah_cost.restype = C.c_double
        L.orc_bvh_sah_cost.argtypes = [vp]
        L.orc_render_f32.restype = C.c_int
        L.orc_render_f32.argtypes = [C.POINTER(Scene), C.POINTER(Settings), fp, fp, fp, fp, C.POINTER(Stats)]
        L.orc_render_f64.restype = C.c_int
        L.orc_render_f64.argtypes = [C.POINTER(Scene), C.POINTER(Settings), C.POINTER(C.c_double), fp, fp, fp,
                                     C.POINTER(Stats)]
        L.orc_camera_ray.argtypes = [C.POINTER(Camera), C.c_float, C.c_float, C.c_float, fp, fp]
        L.orc_display.restype = C.c_float
        L.orc_display.argtypes = [C.c_float, C.c_uint64]
        L.orc_default_threads.restype = C.c_int
        L.orc_texel_index.restype = C.c_uint64
        L.orc_texel_index.argtypes = [C.c_float, C.c_uint32]
        L.orc_get_uv.argtypes = [C.POINTER(Image), C.c_float, C.c_float, fp]
        for name in ("orc_scatter_f32", "orc_scatter_f64"):
            getattr(L, name).argtypes = [C.POINTER(Scene), C.c_uint32, C.c_int32, fp, fp, C.c_float, C.c_float, fp,
                                         fp, fp, C.POINTER(ScatterOut)]
        L.orc_sky_f32.argtypes = [C.POINTER(Scene), fp, fp]
        ip = C.POINTER(C.c_int32)
        L.orc_cast_records_f32.argtypes = [C.POINTER(Scene), C.c_uint32, fp, fp, fp, ip, ip, fp, fp, fp]
        _L = L
    return _L


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def philox(ctr, key) -> np.ndarray:
    """philox4x32_10 over (n, 4) counters and a (2,) key -> (n, 4) uint32."""
    ctr = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32).reshape(-1, 4))
    key = np.ascontiguousarray(np.asarray(key, dtype=np.uint32).reshape(2))
    out = np.zeros_like(ctr)
    lib().orc_philox_batch(_u(ctr), ctr.shape[0], _u(key), _u(out))
    return out


class Bvh:
    """Host-built FP32 binned-SAH BVH of the oracle."""

    def __init__(self, tri: np.ndarray, max_leaf: int = 4):
        self.tri = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 9)
        self.h = lib().orc_bvh_build(_f(self.tri), self.tri.shape[0], max_leaf)

    def validate(self) -> int:
        return lib().orc_bvh_validate(self.h, _f(self.tri), self.tri.shape[0])

    def node_count(self) -> int:
        return lib().orc_bvh_node_count(self.h)

    def sah_cost(self) -> float:
        return lib().orc_bvh_sah_cost(self.h)

    def __del__(self):
        try:
            if self.h:
                lib().orc_bvh_free(self.h)
        except Exception:
            pass


def intersect(tri, org, dirs, mode=BRUTE64, bvh: Bvh | None = None, tnear=1e-5, tfar=np.inf, nthreads=0):
    """Closest hits of SoA rays: returns (prim int32, t, u, v float64); prim -1 = miss."""
    tri = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 9)
    org = np.asarray(org, dtype=np.float32).reshape(-1, 3)
    dirs = np.asarray(dirs, dtype=np.float32).reshape(-1, 3)
    n = org.shape[0]
    cols = [np.ascontiguousarray(org[:, k]) for k in range(3)] + [np.ascontiguousarray(dirs[:, k]) for k in range(3)]
    tn = np.ascontiguousarray(np.broadcast_to(np.float32(tnear), (n,)) if np.isscalar(tnear) else tnear, dtype=np.float32)
    tf = np.ascontiguousarray(np.broadcast_to(np.float32(tfar), (n,)) if np.isscalar(tfar) else tfar, dtype=np.float32)
    prim = np.zeros(n, np.int32)
    t = np.zeros(n, np.float64)
    u = np.zeros(n, np.float64)
    v = np.zeros(n, np.float64)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    lib().orc_intersect_batch(_f(tri), tri.shape[0], bvh.h if bvh is not None else None, mode,
                              *[_f(c) for c in cols], _f(tn), _f(tf), n,
                              prim.ctypes.data_as(C.POINTER(C.c_int32)), dp(t), dp(u), dp(v), nthreads)
    return prim, t, u, v


def _default_sun():
    """entity::sun defaults (components.h:23-29) evaluated in FP32 as the reference does:
    size = pi_f / 48 and direction = glm::normalize(vec3(0.8, -1, 0)) = v * (1 / sqrt(dot(v, v)))."""
    f = np.float32
    v = np.array([0.8, -1.0, 0.0], np.float32)
    dd = f(f(v[0] * v[0]) + f(v[1] * v[1])) + f(v[2] * v[2])
    inv = f(1.0) / np.sqrt(dd)
    return f(3.14159265359) / f(48.0), (v * inv).astype(np.float32)


DEFAULT_SUN_SIZE, DEFAULT_SUN_DIR = _default_sun()


class SceneBinding:
    """Keeps the numpy arrays of an oracle scene alive."""

    def __init__(self, spec, trav=BRUTE64, bvh: Bvh | None = None):
        self.tri = np.ascontiguousarray(spec.corners(), dtype=np.float32)
        uv = spec.corner_uv()
        self.uv = None if uv is None else np.ascontiguousarray(uv, dtype=np.float32)
        self.mi = np.ascontiguousarray(spec.material_indices, dtype=np.uint32)
        self.mats = (Material * max(1, len(spec.materials)))()
        for i, m in enumerate(spec.materials):
            self.mats[i].type = m.type
            self.mats[i].ior = m.ior
            self.mats[i].roughness = m.roughness
            self.mats[i].reflectiveness = m.reflectiveness
            self.mats[i].emission = m.emission
            for k in range(4):
                self.mats[i].colour[k] = m.colour[k]
            self.mats[i].tex = m.tex
        self.textures = [np.ascontiguousarray(t, dtype=np.float32) for t in spec.textures]
        self.timg = (Image * max(1, len(self.textures)))()
        for i, t in enumerate(self.textures):
            self.timg[i] = Image(_f(t), t.shape[1], t.shape[0])
        self.sky = None if spec.sky is None else np.ascontiguousarray(spec.sky, dtype=np.float32)
        self.bvh = bvh
        s = Scene()
        s.tri = _f(self.tri)
        s.uv = _f(self.uv) if self.uv is not None else None
        s.mat_idx = _u(self.mi)
        s.ntri = self.tri.shape[0]
        s.mats = self.mats
        s.nmat = len(spec.materials)
        s.textures = self.timg
        s.ntex = len(self.textures)
        if self.sky is not None:
            s.sky = Image(_f(self.sky), self.sky.shape[1], self.sky.shape[0])
            rot = getattr(spec, "sky_rotation", (0.0, 0.0))
            s.sky_rot[0], s.sky_rot[1] = float(rot[0]), float(rot[1])
        sun = spec.sun or {}
        s.sun.size = sun.get("size", DEFAULT_SUN_SIZE)
        s.sun.intensity = sun.get("intensity", 100.0)
        if "direction" in sun:
            d = np.asarray(sun["direction"], dtype=np.float64)
            d = (d / np.linalg.norm(d)).astype(np.float32)
        else:
            d = DEFAULT_SUN_DIR
        for k in range(3):
            s.sun.dir[k] = d[k]
            s.sun.colour[k] = sun.get("colour", (1.0, 0.9, 0.7))[k]
        s.sun_enabled = 1 if spec.sun_enabled else 0
        cam = spec.camera
        for k in range(3):
            s.cam.pos[k] = cam["position"][k]
        s.cam.fov = cam["fov"]
        m = np.asarray(cam["matrix"], dtype=np.float32).reshape(16)
        for k in range(16):
            s.cam.matrix[k] = m[k]
        s.cam.mode = cam.get("mode", 0)
        s.cam.scale = cam.get("scale", 1.0)
        s.bvh = bvh.h if bvh is not None else None
        s.trav = trav
        self.c = s


def render(spec, res=None, spp=1, max_bounces=None, rr_start=None, mode=None, seed=0, sample_offset=0,
           aspect=None, precision=32, trav=BRUTE64, bvh=None, rows=None, nthreads=0, aovs=True):
    """Runs the oracle integrator; returns dict(raw (H,W,3), albedo/normal/depth (H,W,4), stats)."""
    W, H = res or spec.res
    b = SceneBinding(spec, trav=trav, bvh=bvh)
    st = Settings()
    st.res_x, st.res_y = W, H
    st.max_bounces = spec.max_bounces if max_bounces is None else max_bounces
    st.spp = spp
    st.sample_offset = sample_offset
    st.rr_start = spec.rr_start if rr_start is None else rr_start
    st.mode = spec.mode if mode is None else mode
    st.seed = seed
    st.aspect = (W / H) if aspect is None else aspect
    if rows is not None:
        st.row_begin, st.row_end = rows
    st.nthreads = nthreads
    raw = np.zeros((H, W, 3), np.float64 if precision == 64 else np.float32)
    aov = [np.zeros((H, W, 4), np.float32) for _ in range(3)] if aovs else [None, None, None]
    stats = Stats()
    ap = [(_f(a) if a is not None else None) for a in aov]
    if precision == 64:
        rc = lib().orc_render_f64(C.byref(b.c), C.byref(st), raw.ctypes.data_as(C.POINTER(C.c_double)), *ap,
                                  C.byref(stats))
    else:
        rc = lib().orc_render_f32(C.byref(b.c), C.byref(st), _f(raw), *ap, C.byref(stats))
    if rc != 0:
        raise RuntimeError(f"oracle render failed: {rc}")
    return {"raw": raw, "albedo": aov[0], "normal": aov[1], "depth": aov[2],
            "stats": {"ext_rays": stats.ext_rays, "shadow_rays": stats.shadow_rays,
                      "ref_segments": stats.ref_segments, "seconds": stats.seconds}}


def cast_records(spec, org, dirs, trav=BRUTE64, bvh: Bvh | None = None) -> dict:
    """scene::cast_ray (scene.cpp:79-98) per ray in FP32 through the integrator's own cast:
    dict(distance, prim_id, material (-1 / +inf on a miss), uv (n, 2), normal, point (n, 3))."""
    org = np.ascontiguousarray(org, dtype=np.float32).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
    n = org.shape[0]
    b = SceneBinding(spec, trav=trav, bvh=bvh)
    out = {"distance": np.empty(n, np.float32), "prim_id": np.empty(n, np.int32), "material": np.empty(n, np.int32),
           "uv": np.empty((n, 2), np.float32), "normal": np.empty((n, 3), np.float32),
           "point": np.empty((n, 3), np.float32)}
    ip = C.POINTER(C.c_int32)
    lib().orc_cast_records_f32(C.byref(b.c), n, _f(org), _f(dirs), _f(out["distance"]),
                               out["prim_id"].ctypes.data_as(ip), out["material"].ctypes.data_as(ip),
                               _f(out["uv"]), _f(out["normal"]), _f(out["point"]))
    return out


def display(raw: np.ndarray, n: int) -> np.ndarray:
    """renderer.cpp:371-383 display transform."""
    m = np.clip(raw.astype(np.float32) / np.float32(n), 0.0, 1.0)
    return np.power(m, np.float32(1.0 / 2.2))


def intersect_instanced(models, instances, org, dirs, mode=BRUTE64, nthreads=0):
    """Closest hit over models x instances as the reference computes it
    (scene.cpp:79-98 -> model.cpp:99-126): per instance, inverse(T) maps the ray into object
    space with the direction renormalised; the object-space hit (tnear 1e-5) is mapped back
    by T and its distance recomputed in world space as |T p - o|; the strictly nearer hit
    wins, so the first model / instance wins ties.

    models: list of (ntri, 9) float32 corner arrays; instances: list (per model) of lists of
    column-major 4x4 object->world transforms.  Returns (model, inst, prim, dist, u, v) with
    model = -1 on a miss; dist in float64."""
    org = np.asarray(org, np.float64).reshape(-1, 3)
    dirs = np.asarray(dirs, np.float64).reshape(-1, 3)
    n = org.shape[0]
    best = np.full(n, np.inf)
    bm = np.full(n, -1, np.int64)
    bi = np.full(n, -1, np.int64)
    bp = np.full(n, -1, np.int64)
    bu = np.zeros(n)
    bv = np.zeros(n)
    for k, (tri, insts) in enumerate(zip(models, instances)):
        for j, T in enumerate(insts):
            M = np.asarray(T, np.float64).reshape(4, 4).T        # column-major -> row-major
            inv = np.linalg.inv(M)
            o2 = (org @ inv[:3, :3].T) + inv[:3, 3]
            d2 = dirs @ inv[:3, :3].T
            d2 = d2 / np.linalg.norm(d2, axis=1, keepdims=True)
            prim, t, u, v = intersect(tri, o2.astype(np.float32), d2.astype(np.float32), mode=mode, nthreads=nthreads)
            hit = prim >= 0
            th = np.where(hit, t, 0.0)
            p_obj = o2.astype(np.float32).astype(np.float64) + d2.astype(np.float32).astype(np.float64) * th[:, None]
            p_w = p_obj @ M[:3, :3].T + M[:3, 3]
            dist = np.where(hit, np.linalg.norm(p_w - org, axis=1), np.inf)
            win = dist < best
            best = np.where(win, dist, best)
            bm = np.where(win, k, bm)
            bi = np.where(win, j, bi)
            bp = np.where(win, prim, bp)
            bu = np.where(win, u, bu)
            bv = np.where(win, v, bv)
    return bm, bi, bp, best, bu, bv


def texel_index(u: float, size: int) -> int:
    """Texel column of coordinate u in an image `size` texels wide (image.h:104-109)."""
    return int(lib().orc_texel_index(C.c_float(u), size))


def get_uv(img: np.ndarray, u: float, v: float) -> np.ndarray:
    """image::get_uv of an (h, w, 4) float32 image: the nearest texel with wrap-around."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    out = np.zeros(4, np.float32)
    im = Image(_f(img), img.shape[1], img.shape[0])
    lib().orc_get_uv(C.byref(im), u, v, _f(out))
    return out


def scatter(spec, prim, point, normal, in_d, u=(0, 0, 0, 0), u2=(0, 0, 0, 0), uv=(0.0, 0.0), mode=None,
            precision=32) -> dict:
    """One process_hit step (renderer.cpp:21-102) of the oracle for a hit on triangle `prim`
    of `spec` at `point` with unit normal `normal`, incoming direction in_d and the bounce's
    uniforms u (block 0) and u2 (block 1)."""
    b = SceneBinding(spec)
    f3 = lambda a: np.ascontiguousarray(a, dtype=np.float32).reshape(3)  # noqa: E731
    p, n, d = f3(point), f3(normal), f3(in_d)
    ua = np.ascontiguousarray(u, dtype=np.float32).reshape(4)
    ub = np.ascontiguousarray(u2, dtype=np.float32).reshape(4)
    out = ScatterOut()
    fn = lib().orc_scatter_f64 if precision == 64 else lib().orc_scatter_f32
    fn(C.byref(b.c), spec.mode if mode is None else mode, prim, _f(p), _f(n), uv[0], uv[1], _f(d), _f(ua), _f(ub),
       C.byref(out))
    return {"o": np.array(out.o, np.float32), "d": np.array(out.d, np.float32),
            "albedo": np.array(out.albedo, np.float32), "colour": np.array(out.colour, np.float32),
            "nee_n": np.array(out.nee_n, np.float32), "emission": float(out.emission),
            "alpha": bool(out.flags & SCATTER_ALPHA), "delta": bool(out.flags & SCATTER_DELTA),
            "absorbed": bool(out.flags & SCATTER_ABSORBED)}


def sky(spec, dirs) -> np.ndarray:
    """Radiance of missed rays with directions `dirs` (n, 3): renderer.cpp:277-289 -> scene.cpp:67-77."""
    b = SceneBinding(spec)
    dirs = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
    out = np.zeros_like(dirs)
    for i in range(dirs.shape[0]):
        lib().orc_sky_f32(C.byref(b.c), _f(dirs[i]), _f(out[i]))
    return out
