"""Python mirror of CRender's scene / renderer interface over the crt C ABI.

Method names, argument meaning and error behaviour follow the reference:
  * ``Scene`` mirrors ``cr::scene`` (src/render/scene.h:19-53): add_model, set_skybox,
    set_skybox_rotation, set_sun_enabled, cast_ray (batched here: ``cast_rays``), plus
    set_sun / set_camera for the registry's sun and camera entities.
  * ``Renderer`` mirrors ``cr::renderer`` (src/render/renderer.h:27-68): start, pause,
    update, set_resolution, set_max_bounces, set_target_spp, current_sample_count,
    current_resolution, current_progress / normals / albedos / depths, current_stats.
Failures raise :class:`CrtError` with the library's message; nothing falls back to a
CPU path.
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Callable

import numpy as np

from . import _abi as A
from ._abi import CrtError, check, fptr, uptr, vptr


class Context:
    def __init__(self, device: int = 0, bvh_width: int = 0, max_leaf: int = 0, phase_timing: bool = False,
                 trace_waves_per_cu: int = 0, trace_stack: int = 0, sort_rays: bool = False,
                 sort_materials: bool = False, traversal_stats: bool = False, wavefront_paths: int = 0,
                 trace_refill: int = 0, trace_park_min: int = 0, trace_leaf_rounds: int = 0,
                 bvh_prim_cost: float = 0.0):
        L = A.lib()
        o = A.Options()
        L.crt_options_default(C.byref(o))
        o.device = device
        o.bvh_width = bvh_width
        o.max_leaf = max_leaf
        o.trace_waves_per_cu = trace_waves_per_cu
        o.trace_stack = trace_stack
        o.wavefront_paths = wavefront_paths
        o.trace_refill = trace_refill
        o.trace_park_min = trace_park_min
        o.trace_leaf_rounds = trace_leaf_rounds
        o.bvh_prim_cost = bvh_prim_cost
        o.flags = ((A.CRT_FLAG_PHASE_TIMING if phase_timing else 0) | (A.CRT_FLAG_SORT_RAYS if sort_rays else 0) |
                   (A.CRT_FLAG_SORT_MATERIALS if sort_materials else 0) |
                   (A.CRT_FLAG_TRAVERSAL_STATS if traversal_stats else 0))
        h = C.c_void_p()
        check(L.crt_context_create(C.byref(o), C.byref(h)))
        self._h = h
        self._L = L
        # objects created on the context, closed before it in dependency order
        self._children = {k: weakref.WeakSet() for k in ("renderer", "nrc", "scene")}

    def _adopt(self, kind: str, obj) -> None:
        self._children[kind].add(obj)

    @property
    def handle(self):
        return self._h

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        check(self._L.crt_context_device_name(self._h, buf, 256))
        return buf.value.decode()

    def close(self):
        """Closes the renderers, caches and scenes still open on the context, then the
        context (crt_context_destroy refuses a context with live objects)."""
        if self._h:
            for kind in ("renderer", "nrc", "scene"):
                for obj in list(self._children[kind]):
                    obj.close()
            check(self._L.crt_context_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _material(m) -> A.Material:
    x = A.Material()
    A.lib().crt_material_default(C.byref(x))
    x.type = int(m.type)
    x.ior = float(m.ior)
    x.roughness = float(m.roughness)
    x.reflectiveness = float(m.reflectiveness)
    x.emission = float(m.emission)
    for k in range(4):
        x.colour[k] = float(m.colour[k])
    x.tex = int(m.tex)
    return x


class Scene:
    """cr::scene over crt_scene_*; geometry is committed (device BVH build) by commit()."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._L = A.lib()
        h = C.c_void_p()
        check(self._L.crt_scene_create(ctx.handle, C.byref(h)))
        self._h = h
        ctx._adopt("scene", self)
        self._keep = []

    @property
    def handle(self):
        return self._h

    # scene::add_model (scene.cpp:31-34)
    def add_model(self, spec) -> int:
        verts = np.ascontiguousarray(spec.vertices, dtype=np.float32)
        vidx = np.ascontiguousarray(spec.vertex_indices, dtype=np.uint32)
        midx = np.ascontiguousarray(spec.material_indices, dtype=np.uint32)
        md = A.ModelData()
        md.name = spec.name.encode()
        md.vertices = fptr(verts)
        md.num_vertices = verts.shape[0]
        mats = (A.Material * max(1, len(spec.materials)))(*[_material(m) for m in spec.materials])
        md.materials = mats
        md.num_materials = len(spec.materials)
        keep = [verts, vidx, midx, mats]
        if spec.texture_coords is not None and spec.texture_indices is not None:
            tc = np.ascontiguousarray(spec.texture_coords, dtype=np.float32)
            ti = np.ascontiguousarray(spec.texture_indices, dtype=np.uint32)
            md.texture_coords = fptr(tc)
            md.num_texture_coords = tc.shape[0]
            md.texture_indices = uptr(ti)
            keep += [tc, ti]
        texs = [np.ascontiguousarray(t, dtype=np.float32) for t in spec.textures]
        if texs:
            arr = (A.Image * len(texs))()
            for i, t in enumerate(texs):
                arr[i].rgba = fptr(t)
                arr[i].height, arr[i].width = t.shape[0], t.shape[1]
            md.textures = arr
            md.num_textures = len(texs)
            keep += texs + [arr]
        md.vertex_indices = uptr(vidx)
        md.material_indices = uptr(midx)
        md.num_triangles = midx.shape[0]
        mid = C.c_uint32()
        check(self._L.crt_scene_add_model(self._h, C.byref(md), C.byref(mid)))
        return mid.value

    def set_materials(self, model_id: int, materials) -> None:
        mats = (A.Material * max(1, len(materials)))(*[_material(m) for m in materials])
        check(self._L.crt_scene_set_materials(self._h, model_id, mats, len(materials)))

    def set_instances(self, model_id: int, transforms) -> None:
        t = np.ascontiguousarray(np.asarray(transforms, dtype=np.float32).reshape(-1, 16))
        check(self._L.crt_scene_set_instances(self._h, model_id, fptr(t), t.shape[0]))

    # scene::set_skybox (scene.cpp:36-60)
    def set_skybox(self, rgba) -> None:
        if rgba is None:
            check(self._L.crt_scene_set_skybox(self._h, None))
            return
        img = np.ascontiguousarray(rgba, dtype=np.float32)
        im = A.Image(fptr(img), img.shape[1], img.shape[0])
        check(self._L.crt_scene_set_skybox(self._h, C.byref(im)))

    def set_skybox_rotation(self, x: float, y: float) -> None:
        check(self._L.crt_scene_set_skybox_rotation(self._h, x, y))

    def set_sun_enabled(self, value: bool) -> None:
        check(self._L.crt_scene_set_sun_enabled(self._h, 1 if value else 0))

    def set_sun(self, size=None, intensity=None, direction=None, colour=None) -> None:
        s = A.Sun()
        self._L.crt_sun_default(C.byref(s))
        if size is not None:
            s.size = size
        if intensity is not None:
            s.intensity = intensity
        if direction is not None:
            d = np.asarray(direction, dtype=np.float64)
            d = d / np.linalg.norm(d)
            for k in range(3):
                s.direction[k] = d[k]
        if colour is not None:
            for k in range(3):
                s.colour[k] = colour[k]
        check(self._L.crt_scene_set_sun(self._h, C.byref(s)))

    def set_camera(self, cam: dict) -> None:
        c = A.Camera()
        self._L.crt_camera_default(C.byref(c))
        for k in range(3):
            c.position[k] = cam["position"][k]
            c.rotation[k] = cam.get("rotation", (0, 0, 0))[k]
        c.fov = cam.get("fov", 75.0)
        c.mode = cam.get("mode", 0)
        c.scale = cam.get("scale", 1.0)
        m = np.asarray(cam["matrix"], dtype=np.float32).reshape(16)
        for k in range(16):
            c.matrix[k] = m[k]
        check(self._L.crt_scene_set_camera(self._h, C.byref(c)))

    def commit(self) -> None:
        check(self._L.crt_scene_commit(self._h))

    def bvh_info(self) -> dict:
        bi = A.BvhInfo()
        check(self._L.crt_scene_bvh_info(self._h, C.byref(bi)))
        return {f: getattr(bi, f) for f, _ in A.BvhInfo._fields_}

    def bvh_export(self):
        nb = C.c_uint64(0)
        npr = C.c_uint32(0)
        check(self._L.crt_scene_bvh_export(self._h, None, C.byref(nb), None, C.byref(npr)))
        nodes = np.zeros(nb.value, dtype=np.uint8)
        prims = np.zeros(npr.value, dtype=np.uint32)
        check(self._L.crt_scene_bvh_export(self._h, vptr(nodes), C.byref(nb), vptr(prims), C.byref(npr)))
        return nodes, prims

    # scene::cast_ray (scene.cpp:79-98), batched
    def cast_rays(self, org, dirs, tnear=1e-5, tfar=np.inf, normals=False) -> dict:
        org = np.ascontiguousarray(org, dtype=np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        n = org.shape[0]
        comps = [np.ascontiguousarray(org[:, k]) for k in range(3)] + [np.ascontiguousarray(dirs[:, k]) for k in range(3)]
        tn = np.ascontiguousarray(np.broadcast_to(np.float32(tnear), (n,)) if np.isscalar(tnear) else tnear, dtype=np.float32)
        tf = np.ascontiguousarray(np.broadcast_to(np.float32(tfar), (n,)) if np.isscalar(tfar) else tfar, dtype=np.float32)
        rays = A.Rays(*[vptr(a) for a in comps + [tn, tf]])
        out = {"t": np.empty(n, np.float32), "u": np.empty(n, np.float32), "v": np.empty(n, np.float32),
               "prim_id": np.empty(n, np.uint32), "geom_id": np.empty(n, np.uint32),
               "inst_id": np.empty(n, np.uint32)}
        if normals:
            for k in ("ng_x", "ng_y", "ng_z"):
                out[k] = np.empty(n, np.float32)
        hits = A.Hits(*[vptr(out[k]) if k in out else None for k in
                        ("t", "u", "v", "prim_id", "geom_id", "ng_x", "ng_y", "ng_z", "inst_id")])
        check(self._L.crt_intersect(self._h, C.byref(rays), n, C.byref(hits), A.CRT_MEM_HOST))
        return out

    def cast_ray_records(self, org, dirs, tnear=None, tfar=None) -> dict:
        """scene::cast_ray (scene.cpp:79-98) over a batch through crt_cast_rays: the
        intersection record of each ray (distance, prim_id, geom_id, material, u, v,
        n_x..n_z, p_x..p_z).  tnear / tfar default to the reference's 1e-5 and +inf."""
        org = np.ascontiguousarray(org, dtype=np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        n = org.shape[0]
        comps = [np.ascontiguousarray(org[:, k]) for k in range(3)] + [np.ascontiguousarray(dirs[:, k]) for k in range(3)]
        lim = []
        for v in (tnear, tfar):
            lim.append(None if v is None else np.ascontiguousarray(
                np.broadcast_to(np.float32(v), (n,)) if np.isscalar(v) else v, dtype=np.float32))
        rays = A.Rays(*[vptr(a) for a in comps], *[vptr(a) if a is not None else None for a in lim])
        out = {k: np.empty(n, np.uint32 if k in ("prim_id", "geom_id", "material") else np.float32)
               for k in A.RECORD_FIELDS}
        rec = A.Records(*[vptr(out[k]) for k in A.RECORD_FIELDS])
        check(self._L.crt_cast_rays(self._h, C.byref(rays), n, C.byref(rec), A.CRT_MEM_HOST))
        return out

    def cast_ray_records_device(self, rays: dict, records: dict, n: int) -> None:
        """crt_cast_rays on device memory: `rays` maps org_x .. dir_z (and optionally tnear,
        tfar) and `records` maps the crt_records fields to device addresses (ints; missing
        fields are not written, distance is required)."""
        rk = ("org_x", "org_y", "org_z", "dir_x", "dir_y", "dir_z", "tnear", "tfar")
        r = A.Rays(*[C.c_void_p(rays[k]) if rays.get(k) else None for k in rk])
        rec = A.Records(*[C.c_void_p(records[k]) if records.get(k) else None for k in A.RECORD_FIELDS])
        check(self._L.crt_cast_rays(self._h, C.byref(r), n, C.byref(rec), A.CRT_MEM_DEVICE))

    def cast_rays_device(self, rays: dict, hits: dict, n: int, noalloc: bool = False) -> None:
        """crt_intersect on device memory (or crt_intersect_noalloc): `rays` maps org_x ..
        dir_z, tnear, tfar and `hits` maps t, u, v, prim_id, geom_id, ng_x, ng_y, ng_z, inst_id
        to device addresses (ints; missing hit fields are not written)."""
        rk = ("org_x", "org_y", "org_z", "dir_x", "dir_y", "dir_z", "tnear", "tfar")
        hk = ("t", "u", "v", "prim_id", "geom_id", "ng_x", "ng_y", "ng_z", "inst_id")
        r = A.Rays(*[C.c_void_p(rays[k]) for k in rk])
        h = A.Hits(*[C.c_void_p(hits[k]) if hits.get(k) else None for k in hk])
        if noalloc:
            check(self._L.crt_intersect_noalloc(self._h, C.byref(r), n, C.byref(h)))
        else:
            check(self._L.crt_intersect(self._h, C.byref(r), n, C.byref(h), A.CRT_MEM_DEVICE))

    def occluded(self, org, dirs, tnear=1e-5, tfar=np.inf) -> np.ndarray:
        org = np.ascontiguousarray(org, dtype=np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        n = org.shape[0]
        comps = [np.ascontiguousarray(org[:, k]) for k in range(3)] + [np.ascontiguousarray(dirs[:, k]) for k in range(3)]
        tn = np.ascontiguousarray(np.broadcast_to(np.float32(tnear), (n,)) if np.isscalar(tnear) else tnear, dtype=np.float32)
        tf = np.ascontiguousarray(np.broadcast_to(np.float32(tfar), (n,)) if np.isscalar(tfar) else tfar, dtype=np.float32)
        rays = A.Rays(*[vptr(a) for a in comps + [tn, tf]])
        occ = np.zeros(n, np.uint8)
        check(self._L.crt_occluded(self._h, C.byref(rays), n, vptr(occ), A.CRT_MEM_HOST))
        return occ.astype(bool)

    def close(self):
        if self._h:
            check(self._L.crt_scene_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


IMAGE_FORMATS = {"png": 0, "jpg": 1, "exr": 2, "hdr": 3, "pfm": 4}


def write_image(path, rgba, fmt=None) -> None:
    """crt_image_write of an (H, W, 4) float image; the format defaults to the suffix."""
    img = np.ascontiguousarray(rgba, dtype=np.float32)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError("expected an (H, W, 4) image")
    fmt = (fmt or str(path).rsplit(".", 1)[-1]).lower()
    check(A.lib().crt_image_write(str(path).encode(), IMAGE_FORMATS[fmt], fptr(img), img.shape[1], img.shape[0]))


def post_process(ctx: Context, rgba, bloom=False, threshold=0.7, strength=1.0, gray=False, tonemap=False,
                 tonemap_type=0, exposure=1.0, gamma=2.2, full_coverage=False) -> np.ndarray:
    """crt_post_process of an (H, W, 4) image (cr::post_processor::process)."""
    L = A.lib()
    img = np.ascontiguousarray(rgba, dtype=np.float32)
    o = A.PostOptions()
    L.crt_post_options_default(C.byref(o))
    o.bloom, o.bloom_threshold, o.bloom_strength = int(bloom), threshold, strength
    o.gray_scale, o.tonemapping, o.tonemapping_type = int(gray), int(tonemap), tonemap_type
    o.exposure, o.gamma, o.full_coverage = exposure, gamma, int(full_coverage)
    out = np.empty_like(img)
    check(L.crt_post_process(ctx.handle, C.byref(o), vptr(img), vptr(out), img.shape[1], img.shape[0], A.CRT_MEM_HOST))
    return out


def denoise(ctx: Context, colour, normal, albedo, levels=5, sigma_colour=2.0, sigma_normal=0.2,
            sigma_albedo=0.2) -> np.ndarray:
    """crt_denoise: a-trous filter guided by the normal and albedo images (all (H, W, 4))."""
    L = A.lib()
    c, n, a = (np.ascontiguousarray(x, dtype=np.float32) for x in (colour, normal, albedo))
    o = A.DenoiseOptions(levels, sigma_colour, sigma_normal, sigma_albedo)
    out = np.empty_like(c)
    check(L.crt_denoise(ctx.handle, C.byref(o), vptr(c), vptr(n), vptr(a), vptr(out), c.shape[1], c.shape[0],
                        A.CRT_MEM_HOST))
    return out


class ObjModel:
    """crt_obj_load (asset_loader::load_model, asset_loader.cpp:182-303): a parsed OBJ/MTL
    model.  Arrays are copies; add_to() hands the library's own view to a scene."""

    def __init__(self, path, folder=None):
        L = A.lib()
        h = C.c_void_p()
        check(L.crt_obj_load(str(path).encode(), None if folder is None else str(folder).encode(), C.byref(h)))
        self._h, self._L = h, L
        md = A.ModelData()
        check(L.crt_obj_model(h, C.byref(md)))
        self.name = md.name.decode()
        self.warnings = L.crt_obj_warnings(h).decode()
        nv, nt, ntc = md.num_vertices, md.num_triangles, md.num_texture_coords
        self.vertices = np.ctypeslib.as_array(md.vertices, (nv, 3)).copy() if nv else np.zeros((0, 3), np.float32)
        self.vertex_indices = (np.ctypeslib.as_array(md.vertex_indices, (nt, 3)).copy() if nt
                               else np.zeros((0, 3), np.uint32))
        self.material_indices = (np.ctypeslib.as_array(md.material_indices, (nt,)).copy() if nt
                                 else np.zeros(0, np.uint32))
        self.texture_coords = np.ctypeslib.as_array(md.texture_coords, (ntc, 2)).copy() if ntc else None
        self.texture_indices = (np.ctypeslib.as_array(md.texture_indices, (nt, 3)).copy()
                                if ntc and nt else None)
        self.materials = [{"type": m.type, "colour": tuple(m.colour), "emission": m.emission, "tex": m.tex}
                          for m in (md.materials[i] for i in range(md.num_materials))]
        self.textures = [np.ctypeslib.as_array(md.textures[i].rgba,
                                               (md.textures[i].height, md.textures[i].width, 4)).copy()
                         for i in range(md.num_textures)]

    def add_to(self, scene: "Scene") -> int:
        md = A.ModelData()
        check(self._L.crt_obj_model(self._h, C.byref(md)))
        mid = C.c_uint32()
        check(self._L.crt_scene_add_model(scene.handle, C.byref(md), C.byref(mid)))
        return mid.value

    def close(self):
        if self._h:
            self._L.crt_obj_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scene_from_spec(ctx: Context, spec, commit: bool = True) -> Scene:
    """Loads a SceneSpec into a Scene: model, sky, sun, camera; then commits."""
    sc = Scene(ctx)
    sc.add_model(spec)
    sc.set_skybox(spec.sky)
    rot = getattr(spec, "sky_rotation", (0.0, 0.0))
    if rot[0] or rot[1]:
        sc.set_skybox_rotation(float(rot[0]), float(rot[1]))
    sc.set_sun_enabled(spec.sun_enabled)
    if spec.sun:
        sc.set_sun(**spec.sun)
    sc.set_camera(spec.camera)
    if commit:
        sc.commit()
    return sc


class Renderer:
    """cr::renderer (renderer.h:27-68) over crt_renderer_*."""

    def __init__(self, ctx: Context, scene: Scene, res_x: int, res_y: int, bounces: int):
        self.ctx = ctx
        self.scene = scene
        self.nrc = None
        self._L = A.lib()
        h = C.c_void_p()
        check(self._L.crt_renderer_create(ctx.handle, scene.handle, res_x, res_y, bounces, C.byref(h)))
        self._h = h
        ctx._adopt("renderer", self)

    def start(self) -> None:
        check(self._L.crt_renderer_start(self._h))

    def pause(self) -> None:
        check(self._L.crt_renderer_pause(self._h))

    def update(self, fn: Callable[[], None]) -> None:
        """renderer.cpp:185-192: pause, mutate, start (which clears the accumulation)."""
        self.pause()
        fn()
        self.start()

    def wait(self) -> None:
        check(self._L.crt_renderer_wait(self._h))

    def cancel(self) -> None:
        check(self._L.crt_renderer_cancel(self._h))

    def is_running(self) -> bool:
        v = C.c_int()
        check(self._L.crt_renderer_is_running(self._h, C.byref(v)))
        return bool(v.value)

    def set_resolution(self, x: int, y: int) -> None:
        check(self._L.crt_renderer_set_resolution(self._h, x, y))

    def set_max_bounces(self, n: int) -> None:
        check(self._L.crt_renderer_set_max_bounces(self._h, n))

    def set_target_spp(self, n: int) -> None:
        check(self._L.crt_renderer_set_target_spp(self._h, n))

    def set_options(self, mode: int = 0, rr_start: int = 0, seed: int = 0, sample_offset: int = 0) -> None:
        o = A.RenderOptions(mode, rr_start, seed, sample_offset)
        check(self._L.crt_renderer_set_options(self._h, C.byref(o)))

    def render(self, passes: int) -> None:
        check(self._L.crt_renderer_render(self._h, passes))

    def set_nrc(self, nrc, term_bounce=1, footprint_c=0.01, train_paths=65536, train_suffix=4, train_steps=24,
                train=True, warmup_passes=0) -> None:
        """Render with the neural radiance cache `nrc` (None turns it off)."""
        o = A.NrcRenderOptions()
        self._L.crt_nrc_render_options_default(C.byref(o))
        o.term_bounce, o.footprint_c, o.train_paths = term_bounce, footprint_c, train_paths
        o.train_suffix, o.train_steps, o.train = train_suffix, train_steps, 1 if train else 0
        o.warmup_passes = warmup_passes
        check(self._L.crt_renderer_set_nrc(self._h, nrc.handle if nrc is not None else None, C.byref(o)))
        self.nrc = nrc   # the cache outlives its attachment

    def reset(self) -> None:
        check(self._L.crt_renderer_reset(self._h))

    def current_sample_count(self) -> int:
        v = C.c_uint64()
        check(self._L.crt_renderer_current_sample_count(self._h, C.byref(v)))
        return v.value

    def current_resolution(self) -> tuple:
        x, y = C.c_uint32(), C.c_uint32()
        check(self._L.crt_renderer_current_resolution(self._h, C.byref(x), C.byref(y)))
        return x.value, y.value

    def _read(self, which: str) -> np.ndarray:
        W, H = self.current_resolution()
        ch = 3 if which == "raw" else 4
        out = np.empty((H, W, ch), np.float32)
        check(self._L.crt_renderer_read(self._h, A.BUFFER[which], vptr(out)))
        return out

    def current_progress(self) -> np.ndarray:
        return self._read("progress")

    def current_normals(self) -> np.ndarray:
        return self._read("normals")

    def current_albedos(self) -> np.ndarray:
        return self._read("albedos")

    def current_depths(self) -> np.ndarray:
        return self._read("depths")

    def export(self, path, fmt=None, buffer: str = "progress") -> None:
        """asset_loader::export_framebuffer (asset_loader.cpp:348-377) of a renderer buffer."""
        write_image(path, self._read(buffer), fmt)

    def raw(self) -> np.ndarray:
        """_raw_buffer: (H, W, 3) running sums, flipped in x and y."""
        return self._read("raw")

    def current_stats(self) -> dict:
        s = A.Stats()
        check(self._L.crt_renderer_current_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in A.Stats._fields_}

    def checkpoint(self) -> bytes:
        n = C.c_uint64()
        check(self._L.crt_renderer_checkpoint_size(self._h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(self._L.crt_renderer_checkpoint_save(self._h, buf, n.value))
        return buf.raw

    def restore(self, data: bytes) -> None:
        buf = C.create_string_buffer(data, len(data))
        check(self._L.crt_renderer_checkpoint_load(self._h, buf, len(data)))

    def close(self):
        if self._h:
            self._L.crt_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Nrc:
    """Neural radiance cache (crt_nrc_*): 13-float queries -> RGB, trained with Adam."""

    def __init__(self, ctx: Context, max_batch: int = 65536, seed: int = 1, lr: float = 1e-3):
        self._L = A.lib()
        o = A.NrcOptions()
        self._L.crt_nrc_options_default(C.byref(o))
        o.max_batch = max_batch
        o.seed = seed
        o.lr = lr
        h = C.c_void_p()
        check(self._L.crt_nrc_create(ctx.handle, C.byref(o), C.byref(h)))
        self.ctx = ctx
        self._h = h
        ctx._adopt("nrc", self)
        self.n_params = self._L.crt_nrc_param_count()

    @property
    def handle(self):
        return self._h

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check(self._L.crt_nrc_get_params(self._h, vptr(out)))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        check(self._L.crt_nrc_set_params(self._h, vptr(p)))

    def infer(self, queries) -> np.ndarray:
        q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, 13)
        out = np.empty((q.shape[0], 3), np.float32)
        check(self._L.crt_nrc_infer(self._h, vptr(q), q.shape[0], vptr(out), A.CRT_MEM_HOST))
        return out

    def gradient(self, queries, targets):
        q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, 13)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 3)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        check(self._L.crt_nrc_gradient(self._h, vptr(q), vptr(t), q.shape[0], vptr(g), C.byref(loss), A.CRT_MEM_HOST))
        return loss.value, g

    def train_step(self, queries, targets) -> float:
        q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, 13)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 3)
        loss = C.c_float()
        check(self._L.crt_nrc_train_step(self._h, vptr(q), vptr(t), q.shape[0], C.byref(loss), A.CRT_MEM_HOST))
        return loss.value

    def close(self):
        if self._h:
            check(self._L.crt_nrc_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _job_array(jobs, outs=None):
    n = len(jobs)
    arr = (A.Job * n)()
    for i, j in enumerate(jobs):
        W, H = int(j["width"]), int(j["height"])
        arr[i].scene = j["scene"].handle
        arr[i].width, arr[i].height = W, H
        arr[i].spp, arr[i].max_bounces = int(j["spp"]), int(j["max_bounces"])
        arr[i].options = A.RenderOptions(int(j.get("mode", 0)), int(j.get("rr_start", 0)), int(j.get("seed", 0)),
                                         int(j.get("sample_offset", 0)))
        if outs is not None:
            arr[i].rgb_sum = outs[i].ctypes.data
        arr[i].memory = 0
    return arr


def render_batch(ctx: Context, jobs, max_paths: int = 0, use_graph: bool = True):
    """Renders many jobs together (crt_render_batch).  Each job is a dict with keys scene,
    width, height, spp, max_bounces and optionally mode, rr_start, seed, sample_offset.
    Returns (list of (H, W, 3) running sums, per-job latency in ms, stats dict)."""
    L = A.lib()
    n = len(jobs)
    outs = [np.zeros((int(j["height"]), int(j["width"]), 3), np.float32) for j in jobs]
    arr = _job_array(jobs, outs)
    o = A.BatchOptions()
    L.crt_batch_options_default(C.byref(o))
    o.max_paths, o.use_graph = int(max_paths), 1 if use_graph else 0
    ms = (C.c_double * n)()
    st = A.BatchStats()
    check(L.crt_render_batch(ctx.handle, arr, n, C.byref(o), ms, C.byref(st)))
    return outs, list(ms), {f: getattr(st, f) for f, _ in A.BatchStats._fields_}


def autotune(ctx: Context, jobs, tune_width: bool = True) -> dict:
    """crt_autotune over `jobs` (dicts as for render_batch); returns the chosen shape."""
    L = A.lib()
    arr = _job_array(jobs)
    res = A.TuneResult()
    check(L.crt_autotune(ctx.handle, arr, len(jobs), 1 if tune_width else 0, C.byref(res)))
    return {f: getattr(res, f) for f, _ in A.TuneResult._fields_}
