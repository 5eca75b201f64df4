"""ctypes declarations of include/crt.h (the crt C ABI).

Loading is strict: ``lib()`` raises if libcrt.so is missing, was built without an
export that the header declares, or cannot be opened.  There is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
HEADER = ROOT / "include" / "crt.h"
LIBPATH = PKG / "libcrt.so"

CRT_OK = 0
ERRORS = {
    -1: "CRT_ERR_INVALID_ARG",
    -2: "CRT_ERR_INVALID_MESH",
    -3: "CRT_ERR_HIP",
    -4: "CRT_ERR_OUT_OF_MEMORY",
    -5: "CRT_ERR_STATE",
    -6: "CRT_ERR_UNSUPPORTED",
    -7: "CRT_ERR_CANCELLED",
    -8: "CRT_ERR_CHECKPOINT",
}
CRT_MEM_HOST, CRT_MEM_DEVICE = 0, 1
CRT_METAL, CRT_SMOOTH, CRT_GLASS = 0, 1, 2
CRT_MODE_COMPAT, CRT_MODE_PBR = 0, 1
BUFFER = {"progress": 0, "normals": 1, "albedos": 2, "depths": 3, "raw": 4}
CRT_FLAG_PHASE_TIMING = 1
CRT_FLAG_SORT_RAYS = 2
CRT_FLAG_SORT_MATERIALS = 4
CRT_FLAG_TRAVERSAL_STATS = 8


class CrtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("ior", C.c_float), ("roughness", C.c_float),
                ("reflectiveness", C.c_float), ("emission", C.c_float),
                ("colour", C.c_float * 4), ("tex", C.c_int32)]


class Image(C.Structure):
    _fields_ = [("rgba", C.POINTER(C.c_float)), ("width", C.c_uint32), ("height", C.c_uint32)]


class ModelData(C.Structure):
    _fields_ = [("name", C.c_char_p),
                ("vertices", C.POINTER(C.c_float)), ("num_vertices", C.c_uint32),
                ("texture_coords", C.POINTER(C.c_float)), ("num_texture_coords", C.c_uint32),
                ("materials", C.POINTER(Material)), ("num_materials", C.c_uint32),
                ("textures", C.POINTER(Image)), ("num_textures", C.c_uint32),
                ("vertex_indices", C.POINTER(C.c_uint32)),
                ("texture_indices", C.POINTER(C.c_uint32)),
                ("material_indices", C.POINTER(C.c_uint32)),
                ("num_triangles", C.c_uint32)]


class Sun(C.Structure):
    _fields_ = [("size", C.c_float), ("intensity", C.c_float),
                ("direction", C.c_float * 3), ("colour", C.c_float * 3)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("rotation", C.c_float * 3), ("fov", C.c_float),
                ("mode", C.c_int32), ("scale", C.c_float), ("matrix", C.c_float * 16)]


class Options(C.Structure):
    _fields_ = [("device", C.c_int32), ("bvh_width", C.c_uint32), ("max_leaf", C.c_uint32),
                ("trace_waves_per_cu", C.c_uint32), ("flags", C.c_uint32), ("trace_stack", C.c_uint32),
                ("wavefront_paths", C.c_uint32), ("trace_refill", C.c_uint32), ("trace_park_min", C.c_uint32),
                ("trace_leaf_rounds", C.c_uint32), ("bvh_prim_cost", C.c_float)]


class RenderOptions(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("rr_start", C.c_uint32), ("seed", C.c_uint32),
                ("sample_offset", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("rays_per_second", C.c_uint64), ("samples_per_second", C.c_uint64),
                ("total_rays", C.c_uint64), ("running_time", C.c_double),
                ("samples", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("ms_generate", C.c_double), ("ms_extend", C.c_double), ("ms_shade", C.c_double),
                ("ms_shadow", C.c_double), ("ms_accumulate", C.c_double), ("ms_render", C.c_double),
                ("extend_launches", C.c_uint64), ("ms_nrc_train", C.c_double), ("nrc_loss", C.c_float),
                ("node_visits", C.c_uint64), ("leaf_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("wave_node_steps", C.c_uint64), ("wave_leaf_steps", C.c_uint64),
                ("nrc_detached", C.c_uint32), ("nrc_discarded_passes", C.c_uint32),
                ("node_idle_no_ray", C.c_uint64), ("node_idle_done", C.c_uint64),
                ("node_idle_leaf_wait", C.c_uint64), ("ms_extend_busy", C.c_double)]


class BvhInfo(C.Structure):
    _fields_ = [("width", C.c_uint32), ("node_count", C.c_uint32), ("bvh2_node_count", C.c_uint32),
                ("depth", C.c_uint32), ("max_stack", C.c_uint32), ("sah_cost", C.c_double),
                ("build_ms", C.c_double), ("bytes", C.c_uint64)]


class NrcOptions(C.Structure):
    _fields_ = [("max_batch", C.c_uint32), ("seed", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float)]


class Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_bounces", C.c_uint32), ("options", RenderOptions), ("rgb_sum", C.c_void_p),
                ("memory", C.c_int32)]


class BatchOptions(C.Structure):
    _fields_ = [("max_paths", C.c_uint32), ("use_graph", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("rays_per_second", C.c_double), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("total_rays", C.c_uint64), ("replays", C.c_uint32),
                ("max_paths", C.c_uint32), ("used_graph", C.c_uint32), ("bvh_width", C.c_uint32)]


class TuneResult(C.Structure):
    _fields_ = [("bvh_width", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32), ("trace_stack", C.c_uint32),
                ("max_paths", C.c_uint32), ("rays_per_second", C.c_double), ("trials", C.c_uint32)]


class PostOptions(C.Structure):
    _fields_ = [("bloom", C.c_int32), ("bloom_threshold", C.c_float), ("bloom_strength", C.c_float),
                ("gray_scale", C.c_int32), ("tonemapping", C.c_int32), ("tonemapping_type", C.c_int32),
                ("exposure", C.c_float), ("gamma", C.c_float), ("full_coverage", C.c_int32)]


class DenoiseOptions(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("sigma_colour", C.c_float), ("sigma_normal", C.c_float),
                ("sigma_albedo", C.c_float)]


class NrcRenderOptions(C.Structure):
    _fields_ = [("term_bounce", C.c_uint32), ("footprint_c", C.c_float), ("train_paths", C.c_uint32),
                ("train_suffix", C.c_uint32), ("train_steps", C.c_uint32), ("train", C.c_uint32),
                ("warmup_passes", C.c_uint32)]


class Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("org_x", "org_y", "org_z", "dir_x", "dir_y", "dir_z", "tnear", "tfar")]


class Hits(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("t", "u", "v", "prim_id", "geom_id", "ng_x", "ng_y", "ng_z", "inst_id")]


RECORD_FIELDS = ("distance", "prim_id", "geom_id", "material", "u", "v", "n_x", "n_y", "n_z", "p_x", "p_y", "p_z")


class Records(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in RECORD_FIELDS]


P = C.c_void_p
SIGNATURES = {
    "crt_last_error": (C.c_char_p, []),
    "crt_abi_version": (C.c_int, []),
    "crt_options_default": (None, [C.POINTER(Options)]),
    "crt_context_create": (C.c_int, [C.POINTER(Options), C.POINTER(P)]),
    "crt_context_destroy": (C.c_int, [P]),
    "crt_context_device_name": (C.c_int, [P, C.c_char_p, C.c_size_t]),
    "crt_scene_create": (C.c_int, [P, C.POINTER(P)]),
    "crt_scene_destroy": (C.c_int, [P]),
    "crt_scene_add_model": (C.c_int, [P, C.POINTER(ModelData), C.POINTER(C.c_uint32)]),
    "crt_scene_set_instances": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_float), C.c_uint32]),
    "crt_scene_set_materials": (C.c_int, [P, C.c_uint32, C.POINTER(Material), C.c_uint32]),
    "crt_scene_set_skybox": (C.c_int, [P, C.POINTER(Image)]),
    "crt_scene_set_skybox_rotation": (C.c_int, [P, C.c_float, C.c_float]),
    "crt_scene_set_sun_enabled": (C.c_int, [P, C.c_int]),
    "crt_scene_set_sun": (C.c_int, [P, C.POINTER(Sun)]),
    "crt_scene_set_camera": (C.c_int, [P, C.POINTER(Camera)]),
    "crt_scene_commit": (C.c_int, [P]),
    "crt_scene_bvh_info": (C.c_int, [P, C.POINTER(BvhInfo)]),
    "crt_scene_bvh_export": (C.c_int, [P, P, C.POINTER(C.c_uint64), P, C.POINTER(C.c_uint32)]),
    "crt_intersect": (C.c_int, [P, C.POINTER(Rays), C.c_uint32, C.POINTER(Hits), C.c_int]),
    "crt_intersect_noalloc": (C.c_int, [P, C.POINTER(Rays), C.c_uint32, C.POINTER(Hits)]),
    "crt_cast_rays": (C.c_int, [P, C.POINTER(Rays), C.c_uint32, C.POINTER(Records), C.c_int]),
    "crt_occluded": (C.c_int, [P, C.POINTER(Rays), C.c_uint32, P, C.c_int]),
    "crt_renderer_create": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(P)]),
    "crt_renderer_destroy": (C.c_int, [P]),
    "crt_renderer_start": (C.c_int, [P]),
    "crt_renderer_pause": (C.c_int, [P]),
    "crt_renderer_is_running": (C.c_int, [P, C.POINTER(C.c_int)]),
    "crt_renderer_wait": (C.c_int, [P]),
    "crt_renderer_set_resolution": (C.c_int, [P, C.c_uint32, C.c_uint32]),
    "crt_renderer_set_max_bounces": (C.c_int, [P, C.c_uint32]),
    "crt_renderer_set_target_spp": (C.c_int, [P, C.c_uint64]),
    "crt_renderer_set_options": (C.c_int, [P, C.POINTER(RenderOptions)]),
    "crt_renderer_get_options": (C.c_int, [P, C.POINTER(RenderOptions)]),
    "crt_renderer_render": (C.c_int, [P, C.c_uint32]),
    "crt_renderer_reset": (C.c_int, [P]),
    "crt_renderer_current_sample_count": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "crt_renderer_current_resolution": (C.c_int, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "crt_renderer_read": (C.c_int, [P, C.c_int, P]),
    "crt_renderer_device_buffer": (C.c_int, [P, C.c_int, C.POINTER(P)]),
    "crt_renderer_current_stats": (C.c_int, [P, C.POINTER(Stats)]),
    "crt_renderer_cancel": (C.c_int, [P]),
    "crt_renderer_checkpoint_size": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "crt_renderer_checkpoint_save": (C.c_int, [P, P, C.c_uint64]),
    "crt_renderer_checkpoint_load": (C.c_int, [P, P, C.c_uint64]),
    "crt_nrc_options_default": (None, [C.POINTER(NrcOptions)]),
    "crt_nrc_create": (C.c_int, [P, C.POINTER(NrcOptions), C.POINTER(P)]),
    "crt_nrc_destroy": (C.c_int, [P]),
    "crt_nrc_param_count": (C.c_int, []),
    "crt_nrc_get_params": (C.c_int, [P, P]),
    "crt_nrc_set_params": (C.c_int, [P, P]),
    "crt_nrc_infer": (C.c_int, [P, P, C.c_uint32, P, C.c_int]),
    "crt_nrc_gradient": (C.c_int, [P, P, P, C.c_uint32, P, C.POINTER(C.c_float), C.c_int]),
    "crt_nrc_train_step": (C.c_int, [P, P, P, C.c_uint32, C.POINTER(C.c_float), C.c_int]),
    "crt_nrc_render_options_default": (None, [C.POINTER(NrcRenderOptions)]),
    "crt_renderer_set_nrc": (C.c_int, [P, P, C.POINTER(NrcRenderOptions)]),
    "crt_image_write": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_float), C.c_uint32, C.c_uint32]),
    "crt_post_options_default": (None, [C.POINTER(PostOptions)]),
    "crt_post_process": (C.c_int, [P, C.POINTER(PostOptions), P, P, C.c_uint32, C.c_uint32, C.c_int]),
    "crt_denoise_options_default": (None, [C.POINTER(DenoiseOptions)]),
    "crt_denoise": (C.c_int, [P, C.POINTER(DenoiseOptions), P, P, P, P, C.c_uint32, C.c_uint32, C.c_int]),
    "crt_obj_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "crt_obj_model": (C.c_int, [P, C.POINTER(ModelData)]),
    "crt_obj_warnings": (C.c_char_p, [P]),
    "crt_obj_free": (C.c_int, [P]),
    "crt_batch_options_default": (None, [C.POINTER(BatchOptions)]),
    "crt_autotune": (C.c_int, [P, C.POINTER(Job), C.c_uint32, C.c_uint32, C.POINTER(TuneResult)]),
    "crt_render_batch": (C.c_int, [P, C.POINTER(Job), C.c_uint32, C.POINTER(BatchOptions),
                                   C.POINTER(C.c_double), C.POINTER(BatchStats)]),
    "crt_camera_default": (None, [C.POINTER(Camera)]),
    "crt_camera_update": (None, [C.POINTER(Camera)]),
    "crt_sun_default": (None, [C.POINTER(Sun)]),
    "crt_material_default": (None, [C.POINTER(Material)]),
}


def header_functions() -> list[str]:
    """Names of every function declared in include/crt.h."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(crt_[a-z0-9_]+)\s*\(", text)))


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    """Open libcrt.so (built by crender_amd.build) and bind every declared function."""
    global _LIB
    if _LIB is not None:
        return _LIB
    # CRT_LIBRARY selects another build of the same ABI (e.g. a variant under test)
    path = Path(os.environ.get("CRT_LIBRARY", str(LIBPATH)))
    if not path.exists():
        raise ImportError(f"{path} is missing: run __graft_entry__.build() first")
    L = C.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def check(code: int) -> None:
    if code != CRT_OK:
        raise CrtError(code, lib().crt_last_error().decode(errors="replace"))


def fptr(a):
    """float* of a contiguous float32 numpy array (or None)."""
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def uptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))


def vptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)
